// mqpiv_k2d_wave.hip -- instances and launch of the one-IA-per-wave
// split-precision matrix-core 2-D kernel (mqpiv_wave2d.hpp).
#include "mqpiv_wave2d.hpp"

namespace mqpiv {

namespace {
template <int L, int NTS, int ALG>
hipError_t launch_wave_k(const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs, int64_t n_ia_total,
                         int ns0, int ns1, hipStream_t s) {
  using C = WvCfg<L, NTS>;
  const void* fn = (const void*)piv2d_wave_kernel<L, NTS, ALG>;
  allow_full_lds(fn);
  const size_t lds = C::table_bytes + (size_t)C::WAVES * wv_win_bytes(ns0, ns1);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  // persistent grid: one block (WAVES IAs in flight) per CU
  const int64_t need = (n_ia_total + C::WAVES - 1) / C::WAVES;
  int64_t nb = device_cus();
  if (need < nb) nb = need;
  hipLaunchKernelGGL((piv2d_wave_kernel<L, NTS, ALG>), dim3((unsigned)nb), dim3(64 * C::WAVES), lds, s, j0, d_jobs,
                     d_prefix, n_jobs, n_ia_total, ns0, ns1);
  return hipGetLastError();
}

template <int L, int NTS>
hipError_t launch_wave_l(int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs,
                         int64_t n_ia_total, int ns0, int ns1, hipStream_t s) {
  switch (alg) {
    case ALG_FFTCC: return launch_wave_k<L, NTS, ALG_FFTCC>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, s);
    case ALG_ZNCC: return launch_wave_k<L, NTS, ALG_ZNCC>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, s);
    case ALG_NSQECC: return launch_wave_k<L, NTS, ALG_NSQECC>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, s);
  }
  return hipErrorInvalidValue;
}
}  // namespace

// Transform length of the split-precision form for a 2-D job, or -1 when the
// job is outside what an instance takes (IA size, search size, window, LDS).
int wave2d_len(int alg, int is0, int is1, int ss, int ns0, int ns1) {
  if (alg != ALG_FFTCC && alg != ALG_ZNCC && alg != ALG_NSQECC) return -1;
  using C64 = WvCfg<64, 3>;
  if (ss <= 64 && is0 <= C64::FC && is1 <= C64::FC && ns0 <= 48 && ns1 <= 48 &&
      C64::table_bytes + (size_t)C64::WAVES * wv_win_bytes(ns0, ns1) <= 160 * 1024)
    return 64;
  return -1;
}

size_t wave2d_lds(int L, int ns0, int ns1) {
  if (L == 64) return WvCfg<64, 3>::table_bytes + (size_t)WvCfg<64, 3>::WAVES * wv_win_bytes(ns0, ns1);
  return (size_t)-1;
}

hipError_t launch_wave2d(int L, int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs,
                         int64_t n_ia_total, int ns0, int ns1, hipStream_t s) {
  if (n_ia_total <= 0) return hipSuccess;
  if (L == 64) return launch_wave_l<64, 3>(alg, j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, s);
  return hipErrorInvalidValue;
}

}  // namespace mqpiv
