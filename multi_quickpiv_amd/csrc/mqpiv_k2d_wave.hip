// mqpiv_k2d_wave.hip -- instances and launch of the one-IA-per-wave
// split-precision matrix-core 2-D kernel (mqpiv_wave2d.hpp).
#include <algorithm>

#include "mqpiv_wave2d.hpp"

namespace mqpiv {

namespace {
template <int L, int NTS, int ALG, int NQ>
hipError_t launch_wave_k(const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs, int64_t n_ia_total,
                         int ns0, int ns1, int32_t* refine, hipStream_t s) {
  using C = WvCfg<L, NTS>;
  constexpr int WAVES = WvOcc<NQ>::WAVES;
  const void* fn = (const void*)piv2d_wave_kernel<L, NTS, ALG, NQ>;
  allow_full_lds(fn);
  const size_t lds = C::table_bytes + (size_t)WAVES * wv_win_bytes(ns0, ns1);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  // persistent grid: one block (WAVES IAs in flight) per CU
  const int64_t need = (n_ia_total + WAVES - 1) / WAVES;
  int64_t nb = device_cus();
  if (need < nb) nb = need;
  hipLaunchKernelGGL((piv2d_wave_kernel<L, NTS, ALG, NQ>), dim3((unsigned)nb), dim3(64 * WAVES), lds, s, j0, d_jobs,
                     d_prefix, n_jobs, n_ia_total, ns0, ns1, refine);
  return hipGetLastError();
}

template <int L, int NTS, int NQ>
hipError_t launch_wave_l(int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs,
                         int64_t n_ia_total, int ns0, int ns1, int32_t* refine, hipStream_t s) {
  // only ZNCC flags IAs for the FP64 transforms
  if (refine && alg != ALG_ZNCC) return hipErrorInvalidValue;
  switch (alg) {
    case ALG_FFTCC:
      return launch_wave_k<L, NTS, ALG_FFTCC, NQ>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, nullptr, s);
    case ALG_ZNCC:
      return launch_wave_k<L, NTS, ALG_ZNCC, NQ>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, refine, s);
    case ALG_NSQECC:
      return launch_wave_k<L, NTS, ALG_NSQECC, NQ>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, nullptr, s);
  }
  return hipErrorInvalidValue;
}

// The form's geometry limits: an IA of one sub-block (interSize <= 32) with a
// search size <= 64; or of up to 2 x 2 sub-blocks (interSize <= 64), each of
// which with its margin needs a search part <= 64 (margin <= 16).  Windows
// up to 48 x 48 (three output tiles).
bool wave_one_block(int is0, int is1, int ss) { return is0 <= 32 && is1 <= 32 && ss <= 64; }
bool wave_blocks(int is0, int is1, int sm0, int sm1) { return is0 <= 64 && is1 <= 64 && sm0 <= 16 && sm1 <= 16; }
}  // namespace

size_t wave2d_lds(int L, int ns0, int ns1);
namespace {
}  // namespace

// Transform length of the split-precision form for a 2-D job, or -1 when the
// job is outside what an instance takes (IA size, search size, window, LDS).
// "Length" codes of the form: 64 (one sub-block), 65 (2 x 2 sub-blocks, each at
// transform length 64).
int wave2d_len(int alg, int is0, int is1, int ss, int ns0, int ns1) {
  (void)ss;
  if (alg != ALG_FFTCC && alg != ALG_ZNCC && alg != ALG_NSQECC) return -1;
  if (ns0 > 48 || ns1 > 48) return -1;
  const int sm0 = (ns0 - 1) / 2, sm1 = (ns1 - 1) / 2;
  int code = -1;
  if (wave_one_block(is0, is1, std::max(is0 + 2 * sm0, is1 + 2 * sm1))) code = 64;
  else if (wave_blocks(is0, is1, sm0, sm1)) code = 65;
  if (code < 0 || wave2d_lds(code, ns0, ns1) > 160 * 1024) return -1;
  return code;
}

size_t wave2d_lds(int L, int ns0, int ns1) {
  if (L == 64) return WvCfg<64, 3>::table_bytes + (size_t)WvOcc<1>::WAVES * wv_win_bytes(ns0, ns1);
  if (L == 65) return WvCfg<64, 3>::table_bytes + (size_t)WvOcc<2>::WAVES * wv_win_bytes(ns0, ns1);
  return (size_t)-1;
}

// refine: a refine list (count cleared by refine_reset) that ZNCC appends the
// IAs whose FP32-level results may miss the FP64 results' bar to, or nullptr.
hipError_t launch_wave2d(int L, int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs,
                         int64_t n_ia_total, int ns0, int ns1, int32_t* refine, hipStream_t s) {
  if (n_ia_total <= 0) return hipSuccess;
  if (L == 64) return launch_wave_l<64, 3, 1>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, refine, s);
  if (L == 65) return launch_wave_l<64, 3, 2>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, refine, s);
  return hipErrorInvalidValue;
}

// Clears a refine list's count with a kernel (a launch that may be captured
// into a graph records no memset node).
__global__ void refine_reset_kernel(int32_t* refine) { refine[0] = 0; }

hipError_t refine_reset(int32_t* refine, hipStream_t s) {
  hipLaunchKernelGGL(refine_reset_kernel, dim3(1), dim3(1), 0, s, refine);
  return hipGetLastError();
}

}  // namespace mqpiv
