// mqpiv_means3d.hip -- the sums behind 3-D ZNCC's means, per IA: image 1 over
// the IA and image 2 over its clipped search region (zncc_ops.jl:10, 22), as
// three separable passes instead of two full-volume integral arrays.
//
// Without search-region shifts both boxes of IA (v0, v1, v2) are products of
// per-dimension ranges that depend on v_k alone (piv_ops.jl:38-76), so
//   pass x: for every row (y, z) and IA column v0, the sums over the two
//           x ranges of v0;
//   pass y: for every (v0, v1, z), the sums of the pass-x values over the two
//           y ranges of v1;
//   pass z: for every IA, the sums of the pass-y values over its z ranges.
// Each pass adds its terms in increasing coordinate, in FP64.  On a
// 512 x 512 x 256 volume the passes read each image about once, where the two
// integral arrays wrote and re-read 2 x 541 MB per image.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mqpiv_internal.h"
#include "mqpiv_kernels2d.hpp"

namespace mqpiv {

namespace {

// The IA range [tlf, tlf + is) and the clipped search range [sr, sr + gsz) of
// index v along dim k.
struct Ranges {
  int32_t a0, na, b0, nb;
};
__device__ __forceinline__ Ranges ranges_of(const Job& J, int k, int32_t v) {
  const Geo g = geo_of(v, J.is[k], J.sm[k], J.st[k], (int32_t)J.dims[k]);
  return Ranges{(int32_t)g.tlf, J.is[k], (int32_t)g.sr, g.gsz};
}

template <typename TIN>
__global__ __launch_bounds__(256) void means3d_x_kernel(Job J, double* __restrict__ x1, double* __restrict__ x2) {
  const int64_t nx = J.vf_n[0], rows = J.dims[1] * J.dims[2];
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nx * rows) return;
  const int32_t a = (int32_t)(gid % nx);
  const int64_t row = gid / nx;
  const Ranges r = ranges_of(J, 0, a);
  const gptr<TIN> p1 = as_global<TIN>(J.img1) + row * J.dims[0];
  const gptr<TIN> p2 = as_global<TIN>(J.img2) + row * J.dims[0];
  double s1 = 0.0, s2 = 0.0;
#pragma unroll 8
  for (int32_t i = 0; i < r.na; ++i) s1 += (double)p1[r.a0 + i];
#pragma unroll 8
  for (int32_t i = 0; i < r.nb; ++i) s2 += (double)p2[r.b0 + i];
  x1[gid] = s1;  // [v0 + nx (y + D1 z)]
  x2[gid] = s2;
}

__global__ __launch_bounds__(256) void means3d_y_kernel(Job J, const double* __restrict__ x1,
                                                         const double* __restrict__ x2, double* __restrict__ y1,
                                                         double* __restrict__ y2) {
  const int64_t nx = J.vf_n[0], ny = J.vf_n[1], d1 = J.dims[1];
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nx * ny * J.dims[2]) return;
  const int64_t a = gid % nx, rest = gid / nx;
  const int32_t b = (int32_t)(rest % ny);
  const int64_t z = rest / ny;
  const Ranges r = ranges_of(J, 1, b);
  const double* q1 = x1 + a + nx * d1 * z;
  const double* q2 = x2 + a + nx * d1 * z;
  double s1 = 0.0, s2 = 0.0;
#pragma unroll 8
  for (int32_t i = 0; i < r.na; ++i) s1 += q1[nx * (r.a0 + i)];
#pragma unroll 8
  for (int32_t i = 0; i < r.nb; ++i) s2 += q2[nx * (r.b0 + i)];
  y1[gid] = s1;  // [v0 + nx (v1 + ny z)]
  y2[gid] = s2;
}

__global__ __launch_bounds__(256) void means3d_z_kernel(Job J, const double* __restrict__ y1,
                                                         const double* __restrict__ y2, double* __restrict__ out) {
  const int64_t nx = J.vf_n[0], ny = J.vf_n[1], n = J.n_ia;
  const int64_t ia = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= n) return;
  const int64_t ab = ia % (nx * ny);
  const int32_t c = (int32_t)(ia / (nx * ny));
  const Ranges r = ranges_of(J, 2, c);
  double s1 = 0.0, s2 = 0.0;
#pragma unroll 8
  for (int32_t i = 0; i < r.na; ++i) s1 += y1[ab + nx * ny * (r.a0 + i)];
#pragma unroll 8
  for (int32_t i = 0; i < r.nb; ++i) s2 += y2[ab + nx * ny * (r.b0 + i)];
  out[ia] = s1;      // image 1 over the IA
  out[n + ia] = s2;  // image 2 over the clipped search region
}

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

// Bytes launch_means3d takes: the 2 n_ia sums, then the passes' values.
size_t means3d_bytes(const Job& j) {
  const size_t nx = (size_t)j.vf_n[0], ny = (size_t)j.vf_n[1];
  const size_t xs = nx * (size_t)j.dims[1] * (size_t)j.dims[2], ys = nx * ny * (size_t)j.dims[2];
  return al256(2 * (size_t)j.n_ia * sizeof(double)) + 2 * al256(xs * sizeof(double)) +
         2 * al256(ys * sizeof(double));
}

// The per-IA sums of an unshifted 3-D job into buf (means3d_bytes of it):
// buf[ia] = image 1 over IA ia, buf[n_ia + ia] = image 2 over its clipped
// search region.
hipError_t launch_means3d(const Job& j, void* buf, hipStream_t s) {
  if (j.n_ia <= 0) return hipSuccess;
  if (j.shift) return hipErrorInvalidValue;  // shifted regions are not separable
  const size_t nx = (size_t)j.vf_n[0], ny = (size_t)j.vf_n[1];
  const size_t xs = nx * (size_t)j.dims[1] * (size_t)j.dims[2], ys = nx * ny * (size_t)j.dims[2];
  char* b = (char*)buf;
  double* out = (double*)b;
  b += al256(2 * (size_t)j.n_ia * sizeof(double));
  double* x1 = (double*)b;
  b += al256(xs * sizeof(double));
  double* x2 = (double*)b;
  b += al256(xs * sizeof(double));
  double* y1 = (double*)b;
  b += al256(ys * sizeof(double));
  double* y2 = (double*)b;
  const unsigned bx = (unsigned)((xs + 255) / 256), by = (unsigned)((ys + 255) / 256);
  const unsigned bz = (unsigned)((j.n_ia + 255) / 256);
  if (j.in_dtype == DT_F32)
    hipLaunchKernelGGL(means3d_x_kernel<float>, dim3(bx), dim3(256), 0, s, j, x1, x2);
  else
    hipLaunchKernelGGL(means3d_x_kernel<double>, dim3(bx), dim3(256), 0, s, j, x1, x2);
  hipLaunchKernelGGL(means3d_y_kernel, dim3(by), dim3(256), 0, s, j, x1, x2, y1, y2);
  hipLaunchKernelGGL(means3d_z_kernel, dim3(bz), dim3(256), 0, s, j, y1, y2, out);
  return hipGetLastError();
}

}  // namespace mqpiv
