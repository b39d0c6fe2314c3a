// mqpiv_kernels3d.hip -- host dispatch of the 3-D pipeline over scalar type and
// transform length, chunked over IAs (scratch of at most 2 GiB a chunk).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>

#include "mqpiv_internal.h"

namespace mqpiv {

template <int L, typename T, int LPL>
size_t scratch3_per_ia(int alg, int ns0, int ns1, int ns2);
template <int L, typename T, int LPL>
hipError_t launch3d_chunk(const Job& J, void* scratch, int64_t chunk_cap, int64_t ia_base, int64_t n,
                          const double* sat1_img1, hipStream_t s);
extern template size_t scratch3_per_ia<16, float, 1>(int, int, int, int);
extern template hipError_t launch3d_chunk<16, float, 1>(const Job&, void*, int64_t, int64_t, int64_t,
                                                          const double*, hipStream_t);
extern template size_t scratch3_per_ia<32, float, 2>(int, int, int, int);
extern template hipError_t launch3d_chunk<32, float, 2>(const Job&, void*, int64_t, int64_t, int64_t,
                                                          const double*, hipStream_t);
extern template size_t scratch3_per_ia<48, float, 4>(int, int, int, int);
extern template hipError_t launch3d_chunk<48, float, 4>(const Job&, void*, int64_t, int64_t, int64_t,
                                                          const double*, hipStream_t);
extern template size_t scratch3_per_ia<64, float, 4>(int, int, int, int);
extern template hipError_t launch3d_chunk<64, float, 4>(const Job&, void*, int64_t, int64_t, int64_t,
                                                          const double*, hipStream_t);
extern template size_t scratch3_per_ia<16, double, 2>(int, int, int, int);
extern template hipError_t launch3d_chunk<16, double, 2>(const Job&, void*, int64_t, int64_t, int64_t,
                                                          const double*, hipStream_t);
extern template size_t scratch3_per_ia<32, double, 4>(int, int, int, int);
extern template hipError_t launch3d_chunk<32, double, 4>(const Job&, void*, int64_t, int64_t, int64_t,
                                                          const double*, hipStream_t);
extern template size_t scratch3_per_ia<48, double, 4>(int, int, int, int);
extern template hipError_t launch3d_chunk<48, double, 4>(const Job&, void*, int64_t, int64_t, int64_t,
                                                          const double*, hipStream_t);
extern template size_t scratch3_per_ia<64, double, 4>(int, int, int, int);
extern template hipError_t launch3d_chunk<64, double, 4>(const Job&, void*, int64_t, int64_t, int64_t,
                                                          const double*, hipStream_t);

// Transform length for a 3-D search size (both precisions: up to 64), or -1.
int fft_len_3d(int fp64, int ss) {
  (void)fp64;
  static const int lens[] = {16, 32, 48, 64};
  for (int l : lens)
    if (l >= ss) return l;
  return -1;
}

static size_t per_ia(int fp64, int L, int alg, int ns0, int ns1, int ns2) {
  if (fp64) {
    switch (L) {
      case 16: return scratch3_per_ia<16, double, 2>(alg, ns0, ns1, ns2);
      case 32: return scratch3_per_ia<32, double, 4>(alg, ns0, ns1, ns2);
      case 48: return scratch3_per_ia<48, double, 4>(alg, ns0, ns1, ns2);
      case 64: return scratch3_per_ia<64, double, 4>(alg, ns0, ns1, ns2);
    }
  } else {
    switch (L) {
      case 16: return scratch3_per_ia<16, float, 1>(alg, ns0, ns1, ns2);
      case 32: return scratch3_per_ia<32, float, 2>(alg, ns0, ns1, ns2);
      case 48: return scratch3_per_ia<48, float, 4>(alg, ns0, ns1, ns2);
      case 64: return scratch3_per_ia<64, float, 4>(alg, ns0, ns1, ns2);
    }
  }
  return 0;
}

// IAs per chunk: the scratch of one chunk is capped at 2 GiB
// (MQPIV_3D_CHUNK_BYTES overrides).  Fewer, fuller launches: on a 512x512x256
// volume at IA 32^3 a 2 GiB chunk runs 8% faster than a cache-sized 256 MB one,
// and 4 GiB adds under 1%.  A volume with fewer IAs gets a smaller scratch.
int64_t chunk_3d(int fp64, int L, const Job& j) {
  const int ns0 = 2 * j.sm[0] + 1, ns1 = 2 * j.sm[1] + 1, ns2 = 2 * j.sm[2] + 1;
  size_t budget = (size_t)2 << 30;
  if (const char* e = getenv("MQPIV_3D_CHUNK_BYTES")) budget = (size_t)atoll(e);
  int64_t c = (int64_t)(budget / per_ia(fp64, L, j.alg, ns0, ns1, ns2));
  c = std::max<int64_t>(c, 16);
  return std::min<int64_t>(c, j.n_ia);
}

size_t scratch_3d_bytes(int fp64, int L, const Job& j) {
  const int ns0 = 2 * j.sm[0] + 1, ns1 = 2 * j.sm[1] + 1, ns2 = 2 * j.sm[2] + 1;
  return (size_t)chunk_3d(fp64, L, j) * per_ia(fp64, L, j.alg, ns0, ns1, ns2) + 1024;
}

hipError_t launch_piv3d(int fp64, int L, const Job& j, void* scratch, const double* sat1_img1, hipStream_t s) {
  const int64_t cap = chunk_3d(fp64, L, j);
  for (int64_t b = 0; b < j.n_ia; b += cap) {
    const int64_t n = std::min<int64_t>(cap, j.n_ia - b);
    hipError_t e = hipErrorInvalidValue;
    if (fp64) {
      switch (L) {
      case 16: e = launch3d_chunk<16, double, 2>(j, scratch, cap, b, n, sat1_img1, s); break;
      case 32: e = launch3d_chunk<32, double, 4>(j, scratch, cap, b, n, sat1_img1, s); break;
      case 48: e = launch3d_chunk<48, double, 4>(j, scratch, cap, b, n, sat1_img1, s); break;
      case 64: e = launch3d_chunk<64, double, 4>(j, scratch, cap, b, n, sat1_img1, s); break;
      }
    } else {
      switch (L) {
      case 16: e = launch3d_chunk<16, float, 1>(j, scratch, cap, b, n, sat1_img1, s); break;
      case 32: e = launch3d_chunk<32, float, 2>(j, scratch, cap, b, n, sat1_img1, s); break;
      case 48: e = launch3d_chunk<48, float, 4>(j, scratch, cap, b, n, sat1_img1, s); break;
      case 64: e = launch3d_chunk<64, float, 4>(j, scratch, cap, b, n, sat1_img1, s); break;
      }
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace mqpiv
