// mqpiv_kernels2d.hip -- dispatch of the fused 2-D kernel over scalar type and
// transform length; the kernel instances live in mqpiv_k2d_<type>_<L>.hip (one
// translation unit each so they compile in parallel).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

#include "mqpiv_internal.h"
#include "mqpiv_lpl2d.h"

namespace mqpiv {

template <int L, typename T, int LPL>
size_t lds_bytes_2d(int alg, int ns0, int ns1, int ss0);
template <int L, typename T, int LPL>
hipError_t launch_2d_l(int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs,
                       int64_t n_ia_total, int ns0, int ns1, int ss0, char* gz, hipStream_t s, int32_t* refine_out);
extern template size_t lds_bytes_2d<16, float, Lpl2<float, 16>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<16, float, Lpl2<float, 16>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);
extern template size_t lds_bytes_2d<32, float, Lpl2<float, 32>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<32, float, Lpl2<float, 32>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);
extern template size_t lds_bytes_2d<48, float, Lpl2<float, 48>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<48, float, Lpl2<float, 48>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);
extern template size_t lds_bytes_2d<64, float, Lpl2<float, 64>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<64, float, Lpl2<float, 64>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);
extern template size_t lds_bytes_2d<96, float, Lpl2<float, 96>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<96, float, Lpl2<float, 96>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);
extern template size_t lds_bytes_2d<128, float, Lpl2<float, 128>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<128, float, Lpl2<float, 128>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);
extern template size_t lds_bytes_2d<16, double, Lpl2<double, 16>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<16, double, Lpl2<double, 16>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);
extern template size_t lds_bytes_2d<32, double, Lpl2<double, 32>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<32, double, Lpl2<double, 32>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);
extern template size_t lds_bytes_2d<48, double, Lpl2<double, 48>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<48, double, Lpl2<double, 48>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);
extern template size_t lds_bytes_2d<64, double, Lpl2<double, 64>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<64, double, Lpl2<double, 64>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);
extern template size_t lds_bytes_2d<96, double, Lpl2<double, 96>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<96, double, Lpl2<double, 96>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);
extern template size_t lds_bytes_2d<128, double, Lpl2<double, 128>::v>(int, int, int, int);
extern template hipError_t launch_2d_l<128, double, Lpl2<double, 128>::v>(int, const Job&, const Job*, const int64_t*, int,
                                                    int64_t, int, int, int, char*, hipStream_t, int32_t*);

template <int L, typename T, int LPL>
hipError_t launch_2d_refine(int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs,
                            int64_t n_ia_total, int ns0, int ns1, int ss0, const int32_t* list, char* gz,
                            hipStream_t s);
#define MQPIV_REFINE_EXTERN(L)                                                                                    \
  extern template hipError_t launch_2d_refine<L, double, Lpl2<double, L>::v>(                                     \
      int, const Job&, const Job*, const int64_t*, int, int64_t, int, int, int, const int32_t*, char*, hipStream_t);
MQPIV_REFINE_EXTERN(16)
MQPIV_REFINE_EXTERN(32)
MQPIV_REFINE_EXTERN(48)
MQPIV_REFINE_EXTERN(64)
MQPIV_REFINE_EXTERN(96)
MQPIV_REFINE_EXTERN(128)
#undef MQPIV_REFINE_EXTERN

// The split-precision matrix-core form (precisions 2 and 3 below; mqpiv_k2d_wave.hip).
int wave2d_len(int alg, int is0, int is1, int ss, int ns0, int ns1);
size_t wave2d_lds(int L, int ns0, int ns1);
hipError_t launch_wave2d(int L, int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs,
                         int64_t n_ia_total, int ns0, int ns1, int32_t* refine, hipStream_t s);
hipError_t refine_reset(int32_t* refine, hipStream_t s);

// Transform length for a search size: the smallest supported length >= ss
// (up to 128; FP64 at 128 keeps its transform buffer in global memory), or -1.
int fft_len_2d(int fp64, int ss) {
  static const int f32[] = {16, 32, 48, 64, 96, 128};
  static const int f64[] = {16, 32, 48, 64, 96, 128};
  if (fp64) {
    for (int l : f64)
      if (l >= ss) return l;
  } else {
    for (int l : f32)
      if (l >= ss) return l;
  }
  return -1;
}

// fp64: 0 FP32 radix transforms, 1 FP64 radix transforms, 2 the split-precision
// matrix-core form, 3 that form with FP64 transforms recomputing the IAs it
// flags (ZNCC), 4 FP32 radix transforms with FP64 transforms (of the same
// length) recomputing the IAs they flag (masked NSQECC).  The FP64 kernel's LDS
// of forms 3 and 4 is checked by the caller.
size_t lds_2d(int fp64, int L, int alg, int ns0, int ns1, int ss0) {
  if (fp64 == 4) fp64 = 0;
  if (fp64 == 2 || fp64 == 3) return wave2d_lds(L, ns0, ns1);
  if (fp64) {
    switch (L) {
      case 16: return lds_bytes_2d<16, double, Lpl2<double, 16>::v>(alg, ns0, ns1, ss0);
      case 32: return lds_bytes_2d<32, double, Lpl2<double, 32>::v>(alg, ns0, ns1, ss0);
      case 48: return lds_bytes_2d<48, double, Lpl2<double, 48>::v>(alg, ns0, ns1, ss0);
      case 64: return lds_bytes_2d<64, double, Lpl2<double, 64>::v>(alg, ns0, ns1, ss0);
      case 96: return lds_bytes_2d<96, double, Lpl2<double, 96>::v>(alg, ns0, ns1, ss0);
      case 128: return lds_bytes_2d<128, double, Lpl2<double, 128>::v>(alg, ns0, ns1, ss0);
    }
  } else {
    switch (L) {
      case 16: return lds_bytes_2d<16, float, Lpl2<float, 16>::v>(alg, ns0, ns1, ss0);
      case 32: return lds_bytes_2d<32, float, Lpl2<float, 32>::v>(alg, ns0, ns1, ss0);
      case 48: return lds_bytes_2d<48, float, Lpl2<float, 48>::v>(alg, ns0, ns1, ss0);
      case 64: return lds_bytes_2d<64, float, Lpl2<float, 64>::v>(alg, ns0, ns1, ss0);
      case 96: return lds_bytes_2d<96, float, Lpl2<float, 96>::v>(alg, ns0, ns1, ss0);
      case 128: return lds_bytes_2d<128, float, Lpl2<float, 128>::v>(alg, ns0, ns1, ss0);
    }
  }
  return (size_t)-1;
}

// Bytes of global transform buffers (slot flags, then the buffers) a launch of
// n_ia IAs at (fp64, L) needs, or 0 for lengths whose buffer is in LDS.  The
// flags must be zero when a launch starts; every launch leaves them zero.
// Precisions 3 and 4 need their refine list there too: after the FP64 kernel's
// global buffers (form 4 at length 128), at the start otherwise.
static size_t gz_fp64_bytes(int L, int64_t n_ia) {
  if (!(gz_len<128, double>() && L == 128) || n_ia <= 0) return 0;
  return kGzFlagBytes + (size_t)gz_slots_2d(n_ia) * gz_stride_2d<128, double>();
}
size_t gz_scratch_2d(int fp64, int L, int64_t n_ia) {
  if (n_ia <= 0) return 0;
  if (fp64 == 3) return std::max(refine_bytes(n_ia), kGzFlagBytes);  // the allocation clears the flags' bytes
  if (fp64 == 4) return std::max(gz_fp64_bytes(L, n_ia) + refine_bytes(n_ia), kGzFlagBytes);
  if (fp64 == 1) return gz_fp64_bytes(L, n_ia);
  return 0;
}

// The FP64 kernel over a refine list at length Lr (g64: its global transform
// buffers at length 128).
static hipError_t launch_refine_list(int Lr, int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix,
                                     int n_jobs, int64_t n_ia_total, int ns0, int ns1, int ss0, const int32_t* list,
                                     char* g64, hipStream_t s) {
  switch (Lr) {
    case 16: return launch_2d_refine<16, double, Lpl2<double, 16>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, list, g64, s);
    case 32: return launch_2d_refine<32, double, Lpl2<double, 32>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, list, g64, s);
    case 48: return launch_2d_refine<48, double, Lpl2<double, 48>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, list, g64, s);
    case 64: return launch_2d_refine<64, double, Lpl2<double, 64>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, list, g64, s);
    case 96: return launch_2d_refine<96, double, Lpl2<double, 96>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, list, g64, s);
    case 128: return launch_2d_refine<128, double, Lpl2<double, 128>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, list, g64, s);
  }
  return hipErrorInvalidValue;
}

// Lr: precision 3's FP64 transform length (>= the launch's largest search size).
hipError_t launch_piv2d(int fp64, int L, int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix,
                        int n_jobs, int64_t n_ia_total, int ns0, int ns1, int ss0, void* gz, hipStream_t s, int Lr) {
  if (n_ia_total <= 0) return hipSuccess;
  if (fp64 == 2) return launch_wave2d(L, alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, nullptr, s);
  if (fp64 == 3) {
    if (alg != ALG_ZNCC || !gz) return hipErrorInvalidValue;
    int32_t* list = (int32_t*)gz;
    hipError_t e = refine_reset(list, s);
    if (e == hipSuccess) e = launch_wave2d(L, alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, list, s);
    if (e != hipSuccess) return e;
    return launch_refine_list(Lr, alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, list, nullptr, s);
  }
  if (fp64 == 4) {
    if (alg != ALG_MASK_NSQECC || !gz) return hipErrorInvalidValue;
    char* g64 = gz_fp64_bytes(L, n_ia_total) ? (char*)gz : nullptr;
    int32_t* list = (int32_t*)((char*)gz + gz_fp64_bytes(L, n_ia_total));
    hipError_t e = refine_reset(list, s);
    if (e != hipSuccess) return e;
    switch (L) {
      case 16: e = launch_2d_l<16, float, Lpl2<float, 16>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, nullptr, s, list); break;
      case 32: e = launch_2d_l<32, float, Lpl2<float, 32>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, nullptr, s, list); break;
      case 48: e = launch_2d_l<48, float, Lpl2<float, 48>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, nullptr, s, list); break;
      case 64: e = launch_2d_l<64, float, Lpl2<float, 64>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, nullptr, s, list); break;
      case 96: e = launch_2d_l<96, float, Lpl2<float, 96>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, nullptr, s, list); break;
      case 128: e = launch_2d_l<128, float, Lpl2<float, 128>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, nullptr, s, list); break;
      default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    return launch_refine_list(L, alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, list, g64, s);
  }
  if (fp64) {
    switch (L) {
      case 16: return launch_2d_l<16, double, Lpl2<double, 16>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
      case 32: return launch_2d_l<32, double, Lpl2<double, 32>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
      case 48: return launch_2d_l<48, double, Lpl2<double, 48>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
      case 64: return launch_2d_l<64, double, Lpl2<double, 64>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
      case 96: return launch_2d_l<96, double, Lpl2<double, 96>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
      case 128: return launch_2d_l<128, double, Lpl2<double, 128>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
    }
  } else {
    switch (L) {
      case 16: return launch_2d_l<16, float, Lpl2<float, 16>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
      case 32: return launch_2d_l<32, float, Lpl2<float, 32>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
      case 48: return launch_2d_l<48, float, Lpl2<float, 48>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
      case 64: return launch_2d_l<64, float, Lpl2<float, 64>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
      case 96: return launch_2d_l<96, float, Lpl2<float, 96>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
      case 128: return launch_2d_l<128, float, Lpl2<float, 128>::v>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, (char*)gz, s, nullptr);
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace mqpiv
