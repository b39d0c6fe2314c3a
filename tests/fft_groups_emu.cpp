// CPU check of the lane-group FFTs of mqpiv_fft.hpp (the transforms inside the
// 2-D kernels): every lane of a P-lane group is run in turn, and the cross-lane
// exchanges are emulated by replaying the values recorded on the previous pass
// (pass r sees its partners' true values for the first r exchange stages, so
// nst + 1 passes give the exact group result).  Each transform is compared with
// a direct DFT in long double.  Prints one line per case: "L P type dir err".
// Built with -DEMU_WRONG_PARTNER, every exchange reads the wrong lane, which the
// comparison must expose.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../multi_quickpiv_amd/csrc/mqpiv_fft.hpp"

using namespace mqpiv;

template <typename T, int P, int M>
struct ReplayXchg {
  // sent[stage][lane * M + k] from the previous pass; rec = this pass's record
  const std::vector<std::vector<cpx<T>>>* sent;
  std::vector<std::vector<cpx<T>>>* rec;
  int lane;
  int horizon;
  mutable int calls = 0;
  template <int MASK>
  cpx<T> get(cpx<T> v, int k) const {
    const int stage = calls / M;
    ++calls;
    (*rec)[stage][lane * M + k] = v;
#ifdef EMU_WRONG_PARTNER
    // planted mistake (tests only): read the lane next to the partner (a
    // 2-lane group has no other lane: read its own value)
    constexpr int W = P == 2 ? 0 : (MASK == 1 ? 3 : (MASK ^ 1));
#else
    constexpr int W = MASK;
#endif
    static_assert(W < P, "the exchange stays inside the group");
    if (stage < horizon) return (*sent)[stage][(lane ^ W) * M + k];
    return v;
  }
};

template <int L, int P, typename T, bool FWD>
double check(unsigned seed) {
  constexpr int M = L / P;
  constexpr int SIGN = FWD ? -1 : 1;
  constexpr int nst = P == 2 ? 1 : (P == 4 ? 2 : (P == 8 ? 3 : 4));
  std::mt19937 gen(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::vector<T> xr(L), xi(L);
  for (int i = 0; i < L; ++i) {
    xr[i] = (T)u(gen);
    xi[i] = (T)u(gen);
  }
  std::vector<long double> Xr(L, 0), Xi(L, 0);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < L; ++k)
    for (int n = 0; n < L; ++n) {
      const long double a = SIGN * 2 * pi * (long double)((long)k * n % L) / L;
      Xr[k] += (long double)xr[n] * cosl(a) - (long double)xi[n] * sinl(a);
      Xi[k] += (long double)xr[n] * sinl(a) + (long double)xi[n] * cosl(a);
    }
  // forward (DIT): input strided (n = P*k + q), output chunked (chunk(q)*M + k);
  // inverse (DIF): input chunked, output strided
  auto in_index = [](int q, int k) { return FWD ? P * k + q : chunk_of<P>(q) * M + k; };
  auto out_index = [](int q, int k) { return FWD ? chunk_of<P>(q) * M + k : P * k + q; };
  std::vector<std::vector<cpx<T>>> sent(nst, std::vector<cpx<T>>(P * M)), rec = sent;
  std::vector<std::vector<cpx<T>>> res(P, std::vector<cpx<T>>(M));
  for (int horizon = 0; horizon <= nst; ++horizon) {
    for (int q = 0; q < P; ++q) {
      cpx<T> v[M];
      for (int k = 0; k < M; ++k) v[k] = cpx<T>{xr[in_index(q, k)], xi[in_index(q, k)]};
      ReplayXchg<T, P, M> x{&sent, &rec, q, horizon};
      if constexpr (FWD) fft_group_dit<L, P, SIGN, T>(v, q, x);
      else fft_group_dif<L, P, SIGN, T>(v, q, x);
      for (int k = 0; k < M; ++k) res[q][k] = v[k];
    }
    sent = rec;
  }
  double err = 0;
  for (int q = 0; q < P; ++q)
    for (int k = 0; k < M; ++k) {
      const int n = out_index(q, k);
      err = std::fmax(err, std::fabs((double)((long double)res[q][k].x - Xr[n])));
      err = std::fmax(err, std::fabs((double)((long double)res[q][k].y - Xi[n])));
    }
  return err;
}

template <int L, int P, typename T>
void both(const char* tn) {
  std::printf("%d %d %s fwd %.3e\n", L, P, tn, check<L, P, T, true>(17u + L + P));
  std::printf("%d %d %s inv %.3e\n", L, P, tn, check<L, P, T, false>(29u + L + P));
}

int main() {
  // every (length, lanes per line) the 2-D kernels instantiate, and the doubled
  // and quadrupled groups of their last inverse pass
  both<16, 2, float>("f32");
  both<32, 2, float>("f32");
  both<32, 4, float>("f32");
  both<32, 8, float>("f32");
  both<48, 4, float>("f32");
  both<48, 8, float>("f32");
  both<64, 4, float>("f32");
  both<64, 8, float>("f32");
  both<64, 16, float>("f32");
  both<96, 8, float>("f32");
  both<96, 16, float>("f32");
  both<128, 4, float>("f32");
  both<128, 8, float>("f32");
  both<128, 16, float>("f32");
  both<16, 2, double>("f64");
  both<32, 2, double>("f64");
  both<32, 4, double>("f64");
  both<32, 8, double>("f64");
  both<48, 4, double>("f64");
  both<48, 8, double>("f64");
  both<64, 4, double>("f64");
  both<64, 8, double>("f64");
  both<64, 16, double>("f64");
  both<96, 4, double>("f64");
  both<96, 8, double>("f64");
  both<96, 16, double>("f64");
  both<128, 4, double>("f64");
  both<128, 8, double>("f64");
  both<128, 16, double>("f64");
  return 0;
}
