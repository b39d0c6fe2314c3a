// mqpiv_skip.hip -- per-IA skip flags of the double-mask PIV call (2-D and 3-D): the skip
// filter of skip_inter_region with a mask (src/utils/piv_ops.jl:148-157) run on
// image 1 and the vector-field mask, as PIV_CPU_masked with two masks does
// (src/multi_quickPIV.jl:186); the correlation then uses the IA mask.  One wave
// per IA reads the IA's pixels (dim 0 fastest, coalesced), takes both maxima
// and lane 0 writes the flag.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mqpiv_internal.h"

namespace mqpiv {

namespace {

constexpr int kSkipWaves = 4;  // IAs per block

template <typename TIN>
__global__ __launch_bounds__(64 * kSkipWaves) void skip_flags_kernel(Job j, const void* __restrict__ mask_vf,
                                                                     uint8_t* __restrict__ flags) {
  const int64_t ia = (int64_t)blockIdx.x * kSkipWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (ia >= j.n_ia) return;
  // 2-D jobs carry is[2] = vf_n[2] = 1, so one formula serves both
  const int64_t rest = ia / j.vf_n[0];
  const int64_t v0 = ia - rest * j.vf_n[0], v1 = rest % j.vf_n[1], v2 = rest / j.vf_n[1];
  const int64_t tlf0 = v0 * j.st[0], tlf1 = v1 * j.st[1], tlf2 = v2 * j.st[2];  // piv_ops.jl:38-49, 0-based
  const int is0 = j.is[0], is01 = j.is[0] * j.is[1], npx = is01 * j.is[2];
  const TIN* a = (const TIN*)j.img1;
  const TIN* m = (const TIN*)mask_vf;
  // maximum ignoring NaN (a NaN compares false and never replaces the running value)
  double mx1 = -INFINITY, mx2 = -INFINITY;
  for (int p = lane; p < npx; p += 64) {
    const int pc = p / is01, pr = p - pc * is01;
    const int64_t i = (tlf0 + pr % is0) + j.dims[0] * ((tlf1 + pr / is0) + j.dims[1] * (tlf2 + pc));
    if (j.has_threshold) mx1 = fmax(mx1, (double)a[i]);
    if (j.has_mask_threshold) mx2 = fmax(mx2, (double)m[i]);
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    mx1 = fmax(mx1, __shfl_xor(mx1, d));
    mx2 = fmax(mx2, __shfl_xor(mx2, d));
  }
  if (lane == 0) {
    int skip = 0;
    if (j.has_threshold && mx1 < j.threshold) skip = 1;
    if (j.has_mask_threshold && mx2 < j.mask_threshold) skip = 1;
    flags[ia] = (uint8_t)skip;
  }
}

}  // namespace

// Skip flags (one byte per IA) of a masked job from image 1 and mask_vf.
hipError_t launch_skip_flags(const Job& j, const void* mask_vf, uint8_t* flags, hipStream_t s) {
  if (j.n_ia <= 0) return hipSuccess;
  const int64_t blocks = (j.n_ia + kSkipWaves - 1) / kSkipWaves;
  if (j.in_dtype == DT_F32)
    hipLaunchKernelGGL(skip_flags_kernel<float>, dim3((uint32_t)blocks), dim3(64 * kSkipWaves), 0, s, j, mask_vf,
                       flags);
  else
    hipLaunchKernelGGL(skip_flags_kernel<double>, dim3((uint32_t)blocks), dim3(64 * kSkipWaves), 0, s, j,
                       mask_vf, flags);
  return hipGetLastError();
}

}  // namespace mqpiv
