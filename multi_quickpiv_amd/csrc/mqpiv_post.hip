// mqpiv_post.hip -- vector-field post-processing filters on the device
// (src/postprocessing/postprocessing.jl of the reference).
//
// The fields are small next to the frames they come from (one vector per IA,
// N values each) and every filter is a clipped-box stencil over the vector
// grid, so each kernel is one thread per vector reading its box through the
// caches: HBM traffic is the field in and the result out, and the kernels are
// bound by latency and HBM, not arithmetic.  Arithmetic follows the reference's
// Julia code operation by operation in the field's element type (Float32 or
// Float64): every product and sum rounded on its own (no contraction), the
// same reduction order as the CPU restatement (column-major within a box), and
// IEEE division and square root.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mqpiv_post.h"

#pragma clang fp contract(off)

namespace mqpiv {
namespace post {

__device__ inline void box_of(const VGrid& g, int64_t v, int lo[3], int hi[3]) {
  const int c0 = (int)(v % g.n[0]);
  const int64_t q = v / g.n[0];
  const int c1 = (int)(q % g.n[1]);
  const int c2 = (int)(q / g.n[1]);
  const int c[3] = {c0, c1, c2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = c[d] - g.r[d] < 0 ? 0 : c[d] - g.r[d];
    hi[d] = c[d] + g.r[d] > g.n[d] - 1 ? g.n[d] - 1 : c[d] + g.r[d];
  }
}

__device__ inline int64_t lin(const VGrid& g, int x, int y, int z) {
  return (int64_t)x + (int64_t)g.n[0] * ((int64_t)y + (int64_t)g.n[1] * z);
}

template <typename T>
__device__ inline T mag_of(const T* p, int nd) {
  T s = p[0] * p[0];
  for (int c = 1; c < nd; ++c) s = s + p[c] * p[c];
  return sqrt(s);
}

template <typename T>
__device__ inline T ndot(const T* a, T ma, const T* b, T mb, int nd) {
  T s = (a[0] / ma) * (b[0] / mb);
  for (int c = 1; c < nd; ++c) s = s + (a[c] / ma) * (b[c] / mb);
  return s;
}

// x^n, n >= 0, as Julia's ^(::Float64, ::Integer): plain products up to n = 3,
// compensated power-by-squaring (error-free products via fma) above.
__device__ inline double powi(double x, int n) {
  if (n == 0) return 1.0;
  if (n == 1) return x;
  if (n == 2) return x * x;
  if (n == 3) return x * x * x;
  double y = 1.0, xlo = 0.0, ylo = 0.0;
  while (n > 1) {
    if (n & 1) {
      const double err = fma(y, xlo, x * ylo);
      const double p = x * y;
      ylo = fma(x, y, -p) + err;
      y = p;
    }
    const double err = x * 2 * xlo;
    const double p = x * x;
    xlo = fma(x, x, -p) + err;
    x = p;
    n >>= 1;
  }
  const double err = fma(y, xlo, x * ylo);
  return (isfinite(x) && isfinite(err)) ? fma(x, y, err) : x * y;
}

// Base.isless on non-NaN values: -0.0 before +0.0.
template <typename T>
__device__ inline bool less_tot(T a, T b) {
  return a < b || (a == b && signbit(a) && !signbit(b));
}

template <typename T>
__global__ void magnitudes_kernel(VGrid g, const T* __restrict__ VF, T* __restrict__ M) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= g.nv) return;
  M[v] = mag_of(VF + v * g.nd, g.nd);
}

// Deterministic two-level reduction for STD_masking: pass 0 sums x, pass 1 sums
// (x - mean)^2 with mean from stats[0].  Fixed grid, fixed element-to-thread
// assignment and fixed trees, so the result does not depend on timing.
constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 256;

template <typename T>
__global__ void __launch_bounds__(kRedThreads) stats_partial_kernel(const T* __restrict__ x, int64_t n,
                                                                    const double* __restrict__ stats, int pass,
                                                                    double* __restrict__ partials) {
  __shared__ double red[kRedThreads];
  const double mean = pass ? stats[0] : 0.0;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (int64_t)kRedBlocks * kRedThreads) {
    const double a = (double)x[i];
    if (pass) {
      const double d = a - mean;
      s += d * d;
    } else {
      s += a;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kRedBlocks) stats_final_kernel(const double* __restrict__ partials, int64_t n,
                                                                 int pass, double* __restrict__ stats) {
  __shared__ double red[kRedBlocks];
  red[threadIdx.x] = partials[threadIdx.x];
  __syncthreads();
  for (int w = kRedBlocks / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (pass == 0)
      stats[0] = red[0] / (double)n;
    else
      stats[1] = sqrt(red[0] / (double)(n - 1));
  }
}

template <typename T>
__global__ void std_mask_kernel(const T* __restrict__ x, int64_t n, const double* __restrict__ stats, double nsd,
                                uint8_t* __restrict__ mask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  mask[i] = fabs((double)x[i] - stats[0]) > nsd * stats[1];
}

template <typename T>
__global__ void th_mask_kernel(const T* __restrict__ x, int64_t n, double th, int cmp, uint8_t* __restrict__ mask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = (double)x[i];
  mask[i] = cmp ? (a < th) : (a > th);
}

// Replacement of flagged vectors: zero vector (how 0) or, per component, the
// median of the clipped box in the original field (how 1).  The median is
// selected by rank counting over the box (no storage): the k-th order
// statistic is the value with #(< v) <= k < #(<= v) under Base.isless.
template <typename T>
__global__ void replace_kernel(VGrid g, const T* __restrict__ VF, const uint8_t* __restrict__ mask, int how,
                               T* __restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= g.nv) return;
  const int nd = g.nd;
  if (!mask[v]) {
    for (int c = 0; c < nd; ++c) out[v * nd + c] = VF[v * nd + c];
    return;
  }
  if (how == 0) {
    for (int c = 0; c < nd; ++c) out[v * nd + c] = T(0);
    return;
  }
  int lo[3], hi[3];
  box_of(g, v, lo, hi);
  const int64_t m = (int64_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
  const int64_t k1 = (m & 1) ? m / 2 : m / 2 - 1, k2 = m / 2;
  for (int c = 0; c < nd; ++c) {
    bool has_nan = false;
    T a1 = T(0), a2 = T(0);
    for (int z = lo[2]; z <= hi[2]; ++z)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          const T cand = VF[lin(g, x, y, z) * nd + c];
          if (isnan(cand)) {
            has_nan = true;
            continue;
          }
          int64_t nlt = 0, nle = 0;
          for (int z2 = lo[2]; z2 <= hi[2]; ++z2)
            for (int y2 = lo[1]; y2 <= hi[1]; ++y2)
              for (int x2 = lo[0]; x2 <= hi[0]; ++x2) {
                const T o = VF[lin(g, x2, y2, z2) * nd + c];
                nlt += less_tot(o, cand);
                nle += !less_tot(cand, o);  // o <= cand (used only when no value is NaN)
              }
          if (nlt <= k1 && k1 < nle) a1 = cand;
          if (nlt <= k2 && k2 < nle) a2 = cand;
        }
    T med;
    if (has_nan)
      med = T(NAN);
    else if (m & 1)
      med = a1;
    else
      med = a1 / T(2) + a2 / T(2);
    out[v * nd + c] = med;
  }
}

template <typename T>
__global__ void average_kernel(VGrid g, const T* __restrict__ VF, T* __restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= g.nv) return;
  const int nd = g.nd;
  int lo[3], hi[3];
  box_of(g, v, lo, hi);
  for (int c = 0; c < nd; ++c) {
    T s = T(0);
    int64_t m = 0;
    for (int z = lo[2]; z <= hi[2]; ++z)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          const T a = VF[lin(g, x, y, z) * nd + c];
          s = m ? s + a : a;
          ++m;
        }
    out[v * nd + c] = s / (T)m;
  }
}

template <typename T>
__global__ void collectiveness_kernel(VGrid g, const T* __restrict__ VF, const T* __restrict__ M, double min_dot,
                                      float* __restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= g.nv) return;
  const int nd = g.nd;
  int lo[3], hi[3];
  box_of(g, v, lo, hi);
  T vv[3], mv = M[v];
  for (int c = 0; c < nd; ++c) vv[c] = VF[v * nd + c];
  int64_t ns = 0;
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const int64_t u = lin(g, x, y, z);
        ns += (double)ndot(vv, mv, VF + u * nd, M[u], nd) > min_dot;
      }
  const int64_t nn = (int64_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
  out[v] = (float)((double)ns / (double)nn);
}

template <typename T>
__global__ void simavg_kernel(VGrid g, const T* __restrict__ VF, const T* __restrict__ M, double min_dot,
                              int normalize, int scale, int expo, T* __restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= g.nv) return;
  const int nd = g.nd;
  const T mv = M[v];
  if (mv == T(0)) {
    for (int c = 0; c < nd; ++c) out[v * nd + c] = VF[v * nd + c];
    return;
  }
  int lo[3], hi[3];
  box_of(g, v, lo, hi);
  T vv[3];
  for (int c = 0; c < nd; ++c) vv[c] = VF[v * nd + c];
  double avg[3] = {0.0, 0.0, 0.0};
  int64_t ns = 0;
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const int64_t u = lin(g, x, y, z);
        const T mu = M[u];
        const T* pu = VF + u * nd;
        if ((double)ndot(vv, mv, pu, mu, nd) > min_dot) {
          for (int c = 0; c < nd; ++c) avg[c] += (double)((pu[c] / mu) * mu);
          ++ns;
        }
      }
  const int64_t nn = (int64_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
  double fac = 1.0;
  if (normalize) {
    double q = avg[0] * avg[0];
    for (int c = 1; c < nd; ++c) q += avg[c] * avg[c];
    fac = 1.0 / sqrt(q);
  }
  if (scale) fac *= powi((double)ns / (double)nn, expo);
  for (int c = 0; c < nd; ++c) out[v * nd + c] = (T)(avg[c] * fac);
}

template <typename T>
__global__ void lmm_kernel(VGrid g, const T* __restrict__ VF, const T* __restrict__ M, double ratio,
                           T* __restrict__ out, uint8_t* __restrict__ removed) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= g.nv) return;
  const int nd = g.nd;
  int lo[3], hi[3];
  box_of(g, v, lo, hi);
  T s = T(0);
  int64_t m = 0;
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const T a = M[lin(g, x, y, z)];
        s = m ? s + a : a;
        ++m;
      }
  const T mean_mag = (s - M[v]) / (T)(m - 1);
  const bool keep = (double)(M[v] / mean_mag) < ratio;
  for (int c = 0; c < nd; ++c) {
    const T a = VF[v * nd + c];
    out[v * nd + c] = keep ? a : copysign(T(0), a);
  }
  removed[v] = !keep;
}

// ------------------------------------------------------------------ launchers
constexpr int kThreads = 256;

inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

template <typename T>
hipError_t magnitudes_t(const VGrid& g, const void* VF, void* M, hipStream_t s) {
  if (g.nv == 0) return hipSuccess;
  magnitudes_kernel<T><<<grid_for(g.nv), kThreads, 0, s>>>(g, (const T*)VF, (T*)M);
  return hipGetLastError();
}

}  // namespace post

using namespace post;

// Element type: 0 = float, 1 = double (MQPIV_F32 / MQPIV_F64).
#define MQPIV_PP_DISPATCH(dt, call) ((dt) == 0 ? call(float) : call(double))

hipError_t pp_magnitudes(int dt, const VGrid& g, const void* VF, void* M, hipStream_t s) {
#define CALL(T) magnitudes_t<T>(g, VF, M, s)
  return MQPIV_PP_DISPATCH(dt, CALL);
#undef CALL
}

hipError_t pp_std_mask(int dt, int64_t n, const void* x, double nsd, uint8_t* mask, double* scratch,
                       hipStream_t s) {
  if (n == 0) return hipSuccess;
  double* partials = scratch;
  double* stats = scratch + kRedBlocks;
  for (int pass = 0; pass < 2; ++pass) {
    if (dt == 0)
      stats_partial_kernel<float><<<kRedBlocks, kRedThreads, 0, s>>>((const float*)x, n, stats, pass, partials);
    else
      stats_partial_kernel<double><<<kRedBlocks, kRedThreads, 0, s>>>((const double*)x, n, stats, pass, partials);
    stats_final_kernel<<<1, kRedBlocks, 0, s>>>(partials, n, pass, stats);
  }
  if (dt == 0)
    std_mask_kernel<float><<<grid_for(n), kThreads, 0, s>>>((const float*)x, n, stats, nsd, mask);
  else
    std_mask_kernel<double><<<grid_for(n), kThreads, 0, s>>>((const double*)x, n, stats, nsd, mask);
  return hipGetLastError();
}

size_t pp_std_scratch_bytes() { return (kRedBlocks + 2) * sizeof(double); }

hipError_t pp_th_mask(int dt, int64_t n, const void* x, double th, int cmp, uint8_t* mask, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (dt == 0)
    th_mask_kernel<float><<<grid_for(n), kThreads, 0, s>>>((const float*)x, n, th, cmp, mask);
  else
    th_mask_kernel<double><<<grid_for(n), kThreads, 0, s>>>((const double*)x, n, th, cmp, mask);
  return hipGetLastError();
}

hipError_t pp_replace(int dt, const VGrid& g, const void* VF, const uint8_t* mask, int how, void* out,
                      hipStream_t s) {
  if (g.nv == 0) return hipSuccess;
  if (dt == 0)
    replace_kernel<float><<<grid_for(g.nv), kThreads, 0, s>>>(g, (const float*)VF, mask, how, (float*)out);
  else
    replace_kernel<double><<<grid_for(g.nv), kThreads, 0, s>>>(g, (const double*)VF, mask, how, (double*)out);
  return hipGetLastError();
}

hipError_t pp_average(int dt, const VGrid& g, const void* VF, void* out, hipStream_t s) {
  if (g.nv == 0) return hipSuccess;
  if (dt == 0)
    average_kernel<float><<<grid_for(g.nv), kThreads, 0, s>>>(g, (const float*)VF, (float*)out);
  else
    average_kernel<double><<<grid_for(g.nv), kThreads, 0, s>>>(g, (const double*)VF, (double*)out);
  return hipGetLastError();
}

// M: nv magnitudes of the field's element type, computed here into M first.
hipError_t pp_collectiveness(int dt, const VGrid& g, const void* VF, void* M, double min_dot, float* out,
                             hipStream_t s) {
  if (g.nv == 0) return hipSuccess;
  hipError_t e = pp_magnitudes(dt, g, VF, M, s);
  if (e != hipSuccess) return e;
  if (dt == 0)
    collectiveness_kernel<float><<<grid_for(g.nv), kThreads, 0, s>>>(g, (const float*)VF, (const float*)M,
                                                                    min_dot, out);
  else
    collectiveness_kernel<double><<<grid_for(g.nv), kThreads, 0, s>>>(g, (const double*)VF, (const double*)M,
                                                                     min_dot, out);
  return hipGetLastError();
}

hipError_t pp_simavg(int dt, const VGrid& g, const void* VF, void* M, double min_dot, int normalize, int scale,
                     int expo, void* out, hipStream_t s) {
  if (g.nv == 0) return hipSuccess;
  hipError_t e = pp_magnitudes(dt, g, VF, M, s);
  if (e != hipSuccess) return e;
  if (dt == 0)
    simavg_kernel<float><<<grid_for(g.nv), kThreads, 0, s>>>(g, (const float*)VF, (const float*)M, min_dot,
                                                            normalize, scale, expo, (float*)out);
  else
    simavg_kernel<double><<<grid_for(g.nv), kThreads, 0, s>>>(g, (const double*)VF, (const double*)M, min_dot,
                                                             normalize, scale, expo, (double*)out);
  return hipGetLastError();
}

hipError_t pp_lmm(int dt, const VGrid& g, const void* VF, void* M, double ratio, void* out, uint8_t* removed,
                  hipStream_t s) {
  if (g.nv == 0) return hipSuccess;
  hipError_t e = pp_magnitudes(dt, g, VF, M, s);
  if (e != hipSuccess) return e;
  if (dt == 0)
    lmm_kernel<float><<<grid_for(g.nv), kThreads, 0, s>>>(g, (const float*)VF, (const float*)M, ratio,
                                                         (float*)out, removed);
  else
    lmm_kernel<double><<<grid_for(g.nv), kThreads, 0, s>>>(g, (const double*)VF, (const double*)M, ratio,
                                                          (double*)out, removed);
  return hipGetLastError();
}

}  // namespace mqpiv
