// Multipass: the search-region shifts of one pass from the vector field of the
// previous (coarser) pass.
//
// The reference loops `for scale in multipass:-1:1` (src/multi_quickPIV.jl:57-88)
// with interSize, searchMargin and step scaled by `scale` (src/piv_parameters.jl:
// 85-88), and describes the intent at src/utils/piv_ops.jl:160-173: the
// displacements of a coarse pass offset the search regions of the finer IAs they
// contain, interpolated where IAs overlap.  Its offset hook is the `displacement`
// argument of get_search_coordinates (piv_ops.jl:59-76); the loop never passes
// one.  Here the offset of a fine IA is the coarse field interpolated (bilinear /
// trilinear, in vector-field index space) at the IA's centre, rounded to the
// nearest integer (ties to even, as Julia's round).  A NaN among the coarse
// vectors it reads gives offset 0.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "mqpiv_internal.h"

namespace mqpiv {

struct ShiftGeo {
  int32_t nd;
  int32_t nc[3], isc[3], stc[3];  // coarse pass: vector-field size, interSize, step
  int32_t nf[3], isf[3], stf[3];  // fine pass
};

// Interpolation weights of fine coordinate v along one dim: coarse indices a,
// a + 1 (clamped) and the weight w of a + 1.  The fine IA's centre is
// v*stf + (isf-1)/2 and coarse IA u's centre u*stc + (isc-1)/2 (0-based pixels),
// so its coarse-index position is (2*v*stf + (isf-1) - (isc-1)) / (2*stc),
// clamped to [0, nc-1].
__device__ __forceinline__ void interp_1d(int32_t v, int32_t nc, int32_t isc, int32_t stc, int32_t isf,
                                          int32_t stf, int32_t& a, int32_t& b, double& w) {
#pragma clang fp contract(off)
  const int64_t num = 2 * (int64_t)v * stf + (isf - 1) - (isc - 1);
  double t = (double)num / (double)(2 * (int64_t)stc);
  const double tmax = (double)(nc - 1);
  t = t < 0.0 ? 0.0 : (t > tmax ? tmax : t);
  int32_t i = (int32_t)floor(t);
  const int32_t top = nc >= 2 ? nc - 2 : 0;
  i = i > top ? top : i;
  a = i;
  b = i + 1 < nc ? i + 1 : nc - 1;
  w = t - (double)i;
}

__global__ __launch_bounds__(256) void coarse_to_shift_kernel(const double* __restrict__ vfc, ShiftGeo g,
                                                              int64_t n_fine, int32_t* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_fine) return;
  int32_t ia[3] = {0, 0, 0}, ib[3] = {0, 0, 0};
  double w[3] = {0.0, 0.0, 0.0};
  int64_t r = i;
  for (int k = 0; k < g.nd; ++k) {
    const int32_t v = (int32_t)(r % g.nf[k]);
    r /= g.nf[k];
    interp_1d(v, g.nc[k], g.isc[k], g.stc[k], g.isf[k], g.stf[k], ia[k], ib[k], w[k]);
  }
  const int nb2 = g.nd == 3 ? 2 : 1;
  for (int c = 0; c < 3; ++c) {
    int32_t d = 0;
    if (c < g.nd) {
      double acc = 0.0;
      for (int b2 = 0; b2 < nb2; ++b2)
        for (int b1 = 0; b1 < 2; ++b1)
          for (int b0 = 0; b0 < 2; ++b0) {
            double wt = (b0 ? w[0] : 1.0 - w[0]) * (b1 ? w[1] : 1.0 - w[1]);
            if (g.nd == 3) wt = wt * (b2 ? w[2] : 1.0 - w[2]);
            const int64_t u = (int64_t)(b0 ? ib[0] : ia[0]) +
                              (int64_t)g.nc[0] * ((b1 ? ib[1] : ia[1]) + (int64_t)g.nc[1] * (b2 ? ib[2] : ia[2]));
            acc = acc + wt * vfc[g.nd * u + c];
          }
      if (acc == acc) {
        const double lim = 1073741824.0;  // 2^30: keeps the conversion defined
        acc = acc > lim ? lim : (acc < -lim ? -lim : acc);
        d = (int32_t)rint(acc);
      }
    }
    out[3 * i + c] = d;
  }
}

hipError_t launch_coarse_to_shift(const double* vfc, const int32_t* nc, const int32_t* isc, const int32_t* stc,
                                  const int32_t* nf, const int32_t* isf, const int32_t* stf, int nd,
                                  int64_t n_fine, int32_t* out, hipStream_t s) {
  if (n_fine <= 0) return hipSuccess;
  ShiftGeo g;
  g.nd = nd;
  for (int k = 0; k < 3; ++k) {
    g.nc[k] = nc[k];
    g.isc[k] = isc[k];
    g.stc[k] = stc[k];
    g.nf[k] = nf[k];
    g.isf[k] = isf[k];
    g.stf[k] = stf[k];
  }
  const int64_t blocks = (n_fine + 255) / 256;
  hipLaunchKernelGGL(coarse_to_shift_kernel, dim3((unsigned)blocks), dim3(256), 0, s, vfc, g, n_fine, out);
  return hipGetLastError();
}

}  // namespace mqpiv
