// mqpiv_sat.hip -- FP64 integral arrays (summed-area tables) of an image or
// volume, built with on-device prefix scans.
//
// The reference builds, per IA, the squared integral array of the padded search
// window (src/utils/common_ops.jl:142-180) and reads window sums from it by
// inclusion-exclusion (common_ops.jl:203-245).  Here the tables are built once
// per image: S(j0, j1[, j2]) = sum of v over [0, j0) x [0, j1) [x [0, j2)], with
// v = pixel (power 1) or pixel^2 (power 2), so that any window sum -- including
// the reference's zero-padded, quirk-shifted search windows, after clipping to
// the image -- is 4 (2-D) or 8 (3-D) lookups.  Layout: (d0+1) x (d1+1) [x (d2+1)],
// dim 0 fastest, FP64 throughout.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mqpiv_internal.h"
#include "mqpiv_kernels2d.hpp"

namespace mqpiv {

// Dim-0 pass: one wave per output line (j1, j2); inclusive wave scans over 64
// pixels (DPP row shifts and row broadcasts, the steps of the kernels' group
// reductions, which leave every lane's prefix in place) with a carried total,
// four chunks per round with their loads issued together.  Sa receives v^power; Sb (optional) receives v (power 1) from the
// same pass over the image.  Lines with j1 = 0 or j2 = 0 and the j0 = 0
// entries are the zero border.
__global__ __launch_bounds__(64) void sat_dim0_kernel(const void* __restrict__ img, int dt, int power,
                                                      int64_t d0, int64_t d1, int64_t d2, int nd,
                                                      double* __restrict__ Sa, double* __restrict__ Sb) {
  const int64_t D0 = d0 + 1, D1 = d1 + 1;
  const int64_t line = blockIdx.x;
  const int64_t j1 = line % D1;
  const int64_t j2 = line / D1;
  double* outa = Sa + line * D0;
  double* outb = Sb ? Sb + line * D0 : nullptr;
  const int lane = threadIdx.x;
  const bool zero = (j1 == 0) || (nd == 3 && j2 == 0);
  if (lane == 0) {
    outa[0] = 0.0;
    if (outb) outb[0] = 0.0;
  }
  const int64_t src = (zero ? 0 : ((nd == 3 ? (j2 - 1) * d1 : 0) + (j1 - 1))) * d0;
  constexpr int U = 4;
  double ca = 0.0, cb = 0.0;
  for (int64_t c = 0; c < d0; c += 64 * U) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i0 = c + 64 * u + lane;
      const int64_t ic = i0 < d0 ? i0 : d0 - 1;  // in range; discarded below
      v[u] = dt == DT_F32 ? (double)((const float*)img)[src + ic] : ((const double*)img)[src + ic];
      if (zero || i0 >= d0) v[u] = 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i0 = c + 64 * u + lane;
      double a = power == 2 ? v[u] * v[u] : v[u];
      double b = v[u];
      Group<64>::steps([&](auto mv) {
        a += mv(0.0, a);
        if (outb) b += mv(0.0, b);
      });
      a += ca;
      if (i0 < d0) outa[i0 + 1] = a;
      ca = __shfl(a, 63);
      if (outb) {
        b += cb;
        if (i0 < d0) outb[i0 + 1] = b;
        cb = __shfl(b, 63);
      }
    }
  }
}

// Strided pass along dim d (1 or 2), chunked: every thread scans `chunk`
// consecutive entries of one column in place and records the chunk total.
// blockIdx.y selects the table (S, or S + table_stride) and its totals.
constexpr int kScanBatch = 16;

__global__ __launch_bounds__(256) void sat_chunk_scan_kernel(double* __restrict__ S0, int64_t table_stride,
                                                             int64_t n_pos, int64_t len, int64_t stride,
                                                             int64_t pos_mod, int64_t pos_mul, int chunk, int nch,
                                                             double* __restrict__ tot0) {
  double* S = S0 + blockIdx.y * table_stride;
  double* tot = tot0 + blockIdx.y * (n_pos * nch);
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t p = gid % n_pos;
  const int64_t c = gid / n_pos;
  if (c >= nch) return;
  // positions enumerate every index whose dim-d coordinate is 0
  const int64_t base = (p % pos_mod) + (p / pos_mod) * pos_mul;
  const int64_t j0 = c * chunk;
  const int64_t j1 = j0 + chunk < len ? j0 + chunk : len;
  double acc = 0.0;
  for (int64_t jb = j0; jb < j1; jb += kScanBatch) {
    double v[kScanBatch];
#pragma unroll
    for (int u = 0; u < kScanBatch; ++u) {
      const int64_t j = jb + u < j1 ? jb + u : j1 - 1;
      v[u] = S[base + j * stride];
    }
#pragma unroll
    for (int u = 0; u < kScanBatch; ++u) {
      if (jb + u < j1) {
        acc += v[u];
        S[base + (jb + u) * stride] = acc;
      }
    }
  }
  tot[p * nch + c] = acc;
}

// Strided pass along dim d in one sweep: every thread scans one whole column in
// place, its loads issued kScanBatch at a time.  Used when the columns alone give
// enough threads to fill the device (3-D volumes), halving the table traffic of
// the chunked scan + fix.
__global__ __launch_bounds__(256) void sat_column_scan_kernel(double* __restrict__ S0, int64_t table_stride,
                                                              int64_t n_pos, int64_t len, int64_t stride,
                                                              int64_t pos_mod, int64_t pos_mul) {
  double* S = S0 + blockIdx.y * table_stride;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pos) return;
  const int64_t base = (p % pos_mod) + (p / pos_mod) * pos_mul;
  double acc = 0.0;
  for (int64_t jb = 0; jb < len; jb += kScanBatch) {
    double v[kScanBatch];
#pragma unroll
    for (int u = 0; u < kScanBatch; ++u) {
      const int64_t j = jb + u < len ? jb + u : len - 1;
      v[u] = S[base + j * stride];
    }
#pragma unroll
    for (int u = 0; u < kScanBatch; ++u) {
      if (jb + u < len) {
        acc += v[u];
        S[base + (jb + u) * stride] = acc;
      }
    }
  }
}

// Columns at least this many (over the tables built together): one sweep.
constexpr int64_t kColumnSweepMin = 65536;

// Adds the running total of the preceding chunks to every entry of a chunk.
__global__ __launch_bounds__(256) void sat_chunk_fix_kernel(double* __restrict__ S0, int64_t table_stride,
                                                            int64_t n_pos, int64_t len, int64_t stride,
                                                            int64_t pos_mod, int64_t pos_mul, int chunk, int nch,
                                                            const double* __restrict__ tot0) {
  double* S = S0 + blockIdx.y * table_stride;
  const double* tot = tot0 + blockIdx.y * (n_pos * nch);
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t p = gid % n_pos;
  const int64_t c = gid / n_pos + 1;
  if (c >= nch) return;
  double off = 0.0;
  // (unrolled: the loads of a group are in flight together, in the order of the sum)
#pragma unroll 8
  for (int k = 0; k < c; ++k) off += tot[p * nch + k];
  const int64_t base = (p % pos_mod) + (p / pos_mod) * pos_mul;
  const int64_t j0 = c * chunk;
  const int64_t j1 = j0 + chunk < len ? j0 + chunk : len;
  for (int64_t jb = j0; jb < j1; jb += kScanBatch) {
    double v[kScanBatch];
#pragma unroll
    for (int u = 0; u < kScanBatch; ++u) {
      const int64_t j = jb + u < j1 ? jb + u : j1 - 1;
      v[u] = S[base + j * stride];
    }
#pragma unroll
    for (int u = 0; u < kScanBatch; ++u)
      if (jb + u < j1) S[base + (jb + u) * stride] = v[u] + off;
  }
}

size_t sat_bytes(const int64_t* dims, int nd) {
  size_t n = (size_t)(dims[0] + 1) * (dims[1] + 1) * (nd == 3 ? dims[2] + 1 : 1);
  return n * sizeof(double);
}

// Scratch for the chunk totals of up to two tables built together.
size_t sat_tmp_bytes(const int64_t* dims, int nd) {
  const int chunk = 64;
  size_t worst = 0;
  for (int d = 1; d < nd; ++d) {
    int64_t len = dims[d] + 1;
    int64_t nch = (len + chunk - 1) / chunk;
    int64_t npos = (dims[0] + 1) * (nd == 3 ? (d == 1 ? dims[2] + 1 : dims[1] + 1) : 1);
    size_t b = (size_t)npos * nch * sizeof(double);
    if (b > worst) worst = b;
  }
  return 2 * worst;
}

// Builds Sa for `img` (d0 x d1 [x d2], element type dt) with v^power and, when
// Sb is given, Sb with v, in one pass over the image (the strided passes
// address Sb as Sa + (Sb - Sa), one grid row per table).
static hipError_t build_tables(const void* img, int dt, int power, const int64_t* dims, int nd, double* Sa,
                               double* Sb, double* tmp, hipStream_t s) {
  const int64_t d0 = dims[0], d1 = dims[1], d2 = nd == 3 ? dims[2] : 1;
  const int64_t D0 = d0 + 1, D1 = d1 + 1, D2 = nd == 3 ? d2 + 1 : 1;
  hipLaunchKernelGGL(sat_dim0_kernel, dim3((unsigned)(D1 * D2)), dim3(64), 0, s, img, dt, power, d0, d1, d2, nd,
                     Sa, Sb);
  const int ntab = Sb ? 2 : 1;
  const int64_t table_stride = Sb ? (int64_t)(Sb - Sa) : 0;
  const int chunk = 64;
  for (int d = 1; d < nd; ++d) {
    const int64_t len = d == 1 ? D1 : D2;
    const int64_t stride = d == 1 ? D0 : D0 * D1;
    // positions: d = 1 -> (j0, j2) with base j0 + j2*D0*D1; d = 2 -> (j0, j1) with base j0 + j1*D0
    const int64_t pos_mod = D0;
    const int64_t pos_mul = d == 1 ? D0 * D1 : D0;
    const int64_t n_pos = d == 1 ? D0 * D2 : D0 * D1;
    if (n_pos * ntab >= kColumnSweepMin) {
      hipLaunchKernelGGL(sat_column_scan_kernel, dim3((unsigned)((n_pos + 255) / 256), ntab), dim3(256), 0, s, Sa,
                         table_stride, n_pos, len, stride, pos_mod, pos_mul);
      continue;
    }
    const int nch = (int)((len + chunk - 1) / chunk);
    const int64_t n1 = n_pos * nch;
    hipLaunchKernelGGL(sat_chunk_scan_kernel, dim3((unsigned)((n1 + 255) / 256), ntab), dim3(256), 0, s, Sa,
                       table_stride, n_pos, len, stride, pos_mod, pos_mul, chunk, nch, tmp);
    if (nch > 1) {
      const int64_t n2 = n_pos * (nch - 1);
      hipLaunchKernelGGL(sat_chunk_fix_kernel, dim3((unsigned)((n2 + 255) / 256), ntab), dim3(256), 0, s, Sa,
                         table_stride, n_pos, len, stride, pos_mod, pos_mul, chunk, nch, tmp);
    }
  }
  return hipGetLastError();
}

hipError_t build_sat(const void* img, int dt, int power, const int64_t* dims, int nd, double* S, double* tmp,
                     hipStream_t s) {
  return build_tables(img, dt, power, dims, nd, S, nullptr, tmp, s);
}

// The power-2 table S2 and the power-1 table S1 of one image together.
hipError_t build_sat_pair(const void* img, int dt, const int64_t* dims, int nd, double* S2, double* S1,
                          double* tmp, hipStream_t s) {
  return build_tables(img, dt, 2, dims, nd, S2, S1, tmp, s);
}

}  // namespace mqpiv
