// Split-precision DFT-matrix correlation of one 2-D IA on the matrix cores.
//
// The same computation as correlate_2d (the FFT correlation of fftcc.jl:101-126
// and fft_ops.jl:95-105) for transforms of length L = 64 or 96, written as dense
// products with the L-point DFT matrix W(a, b) = exp(-2 pi i ab / L):
//
//   X = Z W          (along dim 2; Z = A + iB, the packed IA and search window)
//   Z^ = W X         (along dim 1)
//   P = conj(A^) B^  (A^, B^ separated from Z^(k) and Z^(-k); columns k1 <= 32)
//   Y = conj(W) P    (along dim 1, rows s0 of the window)
//   C = Re(Y conj(W)) over the kept columns, each column k1 weighted by its
//       multiplicity in the full spectrum (1 for 0 and L/2, 2 in between)
//
// Every real product runs on v_mfma_f32_16x16x32_f16 (v_mfma_f32_16x16x16f16 for
// the last 16-deep step) with split operands: a value
// v is held as hi = fp16(v) and lo = fp16((v - hi) 2^11), and a product a*b is
// a_hi b_hi + 2^-11 (a_hi b_lo + a_lo b_hi), each sum accumulated in FP32 (the
// lo*lo term, about 2^-22 relative, is dropped).  The lo parts carry the 2^11
// so that they stay in FP16's normal range, as do the operands themselves:
// every stage is scaled by a power of two from its worst-case growth so that
// its largest possible value is below 2^15 (the windows to a maximum magnitude
// in [2^12, 2^13)).  Subnormal FP16 operands would lose their low bits.
//
// A 16x16xK product (K = 16 or 32, G = K/4 values a lane) takes A (16xK) from
// lane l as row l%16, k = G(l/16)..+G-1, B (Kx16) as column l%16, the same k,
// and returns D as row 4(l/16)+v, column l%16 (v < 4).  Operands of the four waves of an IA are staged in LDS as
// FP16 rows of stride L + 8 (one buffer: 37 KB at length 64, so that four IAs
// share a CU; 78 KB at 96); the DFT matrices come from constant tables.
#pragma once

#include <hip/hip_runtime.h>

#include "mqpiv_fft.hpp"

namespace mqpiv {

typedef _Float16 mf_h4 __attribute__((ext_vector_type(4)));
typedef _Float16 mf_h8 __attribute__((ext_vector_type(8)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

// Per transform length L (64 or 96): kept columns k1 <= L/2 padded to whole
// 16-column tiles (K1), FP16 row stride of the staged operands (16-byte rows),
// and the LDS of one IA: one buffer of four staged operands (hi/lo of real and
// imaginary parts).  Each stage keeps its results in registers until every wave
// has read the stage's operands, then stages them over the same buffer.
template <int L>
struct MfCfg {
  static_assert(L == 64 || L == 96, "matrix-core form: lengths 64 and 96");
  static constexpr int K1 = L == 64 ? 48 : 64;
  static constexpr int Ld = L + 8;
  static constexpr int P = L * Ld;  // halfs per staged operand
  static constexpr size_t region = (size_t)4 * P * 2;
  static constexpr int XS = L == 64 ? 6 : 7;  // X' = 2^-XS Z' W
  static constexpr int YS = 7;                // Y' = 2^-YS conj(W) P'
};
template <int L>
constexpr size_t mf_region() { return MfCfg<L>::region; }

// DFT tables as hi/lo FP16 pairs (lo scaled by 2^11).  w[p][a * L + b]: p = 0/1
// Re W hi/lo, 2/3 Im W, 4/5 -Im W (W is symmetric, so row a serves as row or
// column).  c[p][s1 * K1 + k1]: p = 0/1 m(k1) Re W(k1, s1), 2/3 m(k1) Im W(k1, s1),
// with the multiplicity m = 1 for k1 = 0 and L/2, 2 for 1..L/2-1 and 0 for the
// padding.
template <int L>
struct alignas(16) MfTables {
  static constexpr int K1 = MfCfg<L>::K1;
  _Float16 w[6][L * L];
  _Float16 c[4][L * K1];
  static constexpr void split(double v, _Float16& hi, _Float16& lo) {
    hi = (_Float16)v;
    lo = (_Float16)((v - (double)hi) * 2048.0);
  }
  constexpr MfTables() : w{}, c{} {
    constexpr Twiddles<L, double> tw{};
    for (int a = 0; a < L; ++a)
      for (int b = 0; b < L; ++b) {
        const int k = (a * b) % L;
        const double re = tw.c[k], im = -tw.s[k];
        const int i = a * L + b;
        split(re, w[0][i], w[1][i]);
        split(im, w[2][i], w[3][i]);
        split(-im, w[4][i], w[5][i]);
      }
    for (int s1 = 0; s1 < L; ++s1)
      for (int k1 = 0; k1 < K1; ++k1) {
        const int k = (k1 * s1) % L;
        const double m = (k1 == 0 || k1 == L / 2) ? 1.0 : (k1 < L / 2 ? 2.0 : 0.0);
        const int i = s1 * K1 + k1;
        split(m * tw.c[k], c[0][i], c[1][i]);
        split(-m * tw.s[k], c[2][i], c[3][i]);
      }
  }
};
static __device__ const MfTables<64> kMfTab64{};
static __device__ const MfTables<96> kMfTab96{};
template <int L>
__device__ __forceinline__ const MfTables<L>& mf_tab() {
  if constexpr (L == 64) return kMfTab64;
  else return kMfTab96;
}

__device__ __forceinline__ mf_f4 mf_mul(mf_h4 a, mf_h4 b, mf_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}
// a * b with split operands: hi*hi into c, the cross terms (carrying 2^11) into x
__device__ __forceinline__ void mf_split(mf_h4 ah, mf_h4 al, mf_h4 bh, mf_h4 bl, mf_f4& c, mf_f4& x) {
  c = mf_mul(ah, bh, c);
  x = mf_mul(ah, bl, x);
  x = mf_mul(al, bh, x);
}
__device__ __forceinline__ float mf_sum(float c, float x) { return c + x * (1.0f / 2048.0f); }
__device__ __forceinline__ void mf_split1(float v, _Float16& hi, _Float16& lo) {
  hi = (_Float16)v;
  lo = (_Float16)((v - (float)hi) * 2048.0f);
}
__device__ __forceinline__ mf_h4 mf_ld(const _Float16* p) { return *reinterpret_cast<const mf_h4*>(p); }
__device__ __forceinline__ mf_h8 mf_ld8(const _Float16* p) { return *reinterpret_cast<const mf_h8*>(p); }
// the same with 16x16x32 products (A: row l%16, k = 8(l/16)..+7; B: column l%16, same k)
__device__ __forceinline__ void mf_split8(mf_h8 ah, mf_h8 al, mf_h8 bh, mf_h8 bl, mf_f4& c, mf_f4& x) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, c, 0, 0, 0);
  x = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, x, 0, 0, 0);
  x = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, x, 0, 0, 0);
}

// Entry and exit as correlate_2d<L, LPL, float>: on entry lane t (line = t/LPL,
// q = t%LPL) holds x[k] = Z(line, LPL k + q); on exit the window C(s0, s1),
// s0 < ns0, s1 < ns1, is at the start of `region` as float, column-major.  One
// IA per block of TPI threads: the L/16 row blocks of every product are spread
// over the waves, and the waves of one row block take its column tiles in turn;
// `red` holds the block reductions.
//
// Scales (powers of two, exact): Z' = s Z with max |Z'| < 2^13 per window;
// X' = 2^-XS Z'W (< 2^14); Z^' = W X'; P' = 2^-28 conj(A^') B^' (< 2^14);
// Y' = 2^-7 conj(W) P' (< 2^14).  So C = C' 2^(2 XS + 35) / (sa sb L^2).
template <int L, int LPL, int TPI, typename Grp>
__device__ __forceinline__ void correlate_2d_mfma(cpx<float> (&x)[L / LPL], void* region, int t, int ns0,
                                                  int ns1, double* red) {
  using Cf = MfCfg<L>;
  constexpr int P = Cf::P, Ld = Cf::Ld, K1 = Cf::K1;
  constexpr int NRB = L / 16;      // row blocks (and column tiles of the forward products)
  constexpr int NW = TPI / 64;     // waves
  static_assert(NW % NRB == 0, "waves per row block");
  constexpr int WPB = NW / NRB;    // waves per row block
  constexpr int KS = L / 32;       // 32-deep steps over a length-L contraction
  const MfTables<L>& tab = mf_tab<L>();
  _Float16* S0 = reinterpret_cast<_Float16*>(region);
  const int line = t / LPL, q = t % LPL;
  // row block and first column tile (with one wave a row block: the wave, and 0)
  const int wv = t >> 6, rb = WPB == 1 ? wv : wv % NRB, cw = WPB == 1 ? 0 : wv / NRB;
  const int l = t & 63, fr = l & 15, fk = (l >> 4) * 4;  // fragment row / column and k offset
  const int fk8 = (l >> 4) * 8;                          // k offset of 32-deep fragments
  const mf_f4 zero4 = {0.f, 0.f, 0.f, 0.f};

  // --- window scales: max |A| and max |B| to [2^12, 2^13)
  double ma = 0.0, mb = 0.0;
#pragma unroll
  for (int k = 0; k < L / LPL; ++k) {
    ma = fmax(ma, (double)fabsf(x[k].x));
    mb = fmax(mb, (double)fabsf(x[k].y));
  }
  Grp::reduce2(ma, mb, OpMax{}, OpMax{}, red, t);
  int ea = 0, eb = 0;
  frexp(ma, &ea);
  frexp(mb, &eb);
  const float sa = ldexpf(1.0f, 13 - ea), sb = ldexpf(1.0f, 13 - eb);

  // --- stage Z row-major [k0][k1]
  if (line < L) {
#pragma unroll
    for (int k = 0; k < L / LPL; ++k) {
      const int i = line * Ld + LPL * k + q;
      mf_split1(x[k].x * sa, S0[i], S0[P + i]);
      mf_split1(x[k].y * sb, S0[2 * P + i], S0[3 * P + i]);
    }
  }
  __syncthreads();

  const _Float16* Wrh = tab.w[0];
  const _Float16* Wrl = tab.w[1];
  const _Float16* Wih = tab.w[2];
  const _Float16* Wil = tab.w[3];
  const _Float16* Wnh = tab.w[4];
  const _Float16* Wnl = tab.w[5];
  constexpr float xsc = 1.0f / (float)(1 << Cf::XS);

  // --- X' = 2^-XS Z' W: rows k0 of block rb; Xr = Zr Wr - Zi Wi, Xi = Zr Wi + Zi Wr.
  // The block's rows of Z go to registers first; X is staged column-major
  // (XT[k1][k0]) over them.
  {
    mf_h8 za[KS][4];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int ai = (16 * rb + fr) * Ld + ks * 32 + fk8;
#pragma unroll
      for (int pp = 0; pp < 4; ++pp) za[ks][pp] = mf_ld8(S0 + pp * P + ai);
    }
    __syncthreads();
#pragma unroll 1
    for (int c = cw; c < NRB; c += WPB) {
      mf_f4 xr = zero4, xr2 = zero4, xi = zero4, xi2 = zero4;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int bi = (16 * c + fr) * L + ks * 32 + fk8;
        const mf_h8 wrh = mf_ld8(Wrh + bi), wrl = mf_ld8(Wrl + bi);
        const mf_h8 wih = mf_ld8(Wih + bi), wil = mf_ld8(Wil + bi);
        const mf_h8 wnh = mf_ld8(Wnh + bi), wnl = mf_ld8(Wnl + bi);
        mf_split8(za[ks][0], za[ks][1], wrh, wrl, xr, xr2);
        mf_split8(za[ks][2], za[ks][3], wnh, wnl, xr, xr2);
        mf_split8(za[ks][0], za[ks][1], wih, wil, xi, xi2);
        mf_split8(za[ks][2], za[ks][3], wrh, wrl, xi, xi2);
      }
      // this lane holds k0 = 16 rb + fk + v, k1 = 16 c + fr
      mf_h4 rh, rl, ih, il;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        _Float16 a, b;
        mf_split1(mf_sum(xr[v], xr2[v]) * xsc, a, b);
        rh[v] = a;
        rl[v] = b;
        mf_split1(mf_sum(xi[v], xi2[v]) * xsc, a, b);
        ih[v] = a;
        il[v] = b;
      }
      const int o = (16 * c + fr) * Ld + 16 * rb + fk;
      *reinterpret_cast<mf_h4*>(S0 + o) = rh;
      *reinterpret_cast<mf_h4*>(S0 + P + o) = rl;
      *reinterpret_cast<mf_h4*>(S0 + 2 * P + o) = ih;
      *reinterpret_cast<mf_h4*>(S0 + 3 * P + o) = il;
    }
  }
  __syncthreads();

  // --- Z^' = W X': rows k0 of block rb; Z^r = Wr Xr - Wi Xi, Z^i = Wi Xr + Wr Xi.  Held in
  // registers until every wave has read X, then written as FP32 complex (row stride L + 1).
  cpx<float>* Zh = reinterpret_cast<cpx<float>*>(region);
  constexpr int NCW = (NRB + WPB - 1) / WPB;  // column tiles per wave
  mf_f4 zhr[NCW], zhi[NCW];
  {
    mf_h8 wa[KS][6];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int ai = (16 * rb + fr) * L + ks * 32 + fk8;
      wa[ks][0] = mf_ld8(Wrh + ai);
      wa[ks][1] = mf_ld8(Wrl + ai);
      wa[ks][2] = mf_ld8(Wih + ai);
      wa[ks][3] = mf_ld8(Wil + ai);
      wa[ks][4] = mf_ld8(Wnh + ai);
      wa[ks][5] = mf_ld8(Wnl + ai);
    }
#pragma unroll
    for (int j = 0; j < NCW; ++j) {
      const int c = cw + WPB * j;
      mf_f4 zr = zero4, zr2 = zero4, zi = zero4, zi2 = zero4;
      if (c < NRB) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int bi = (16 * c + fr) * Ld + ks * 32 + fk8;
          const mf_h8 xrh = mf_ld8(S0 + bi), xrl = mf_ld8(S0 + P + bi);
          const mf_h8 xih = mf_ld8(S0 + 2 * P + bi), xil = mf_ld8(S0 + 3 * P + bi);
          mf_split8(wa[ks][0], wa[ks][1], xrh, xrl, zr, zr2);
          mf_split8(wa[ks][4], wa[ks][5], xih, xil, zr, zr2);
          mf_split8(wa[ks][2], wa[ks][3], xrh, xrl, zi, zi2);
          mf_split8(wa[ks][0], wa[ks][1], xih, xil, zi, zi2);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        zhr[j][v] = mf_sum(zr[v], zr2[v]);
        zhi[j][v] = mf_sum(zi[v], zi2[v]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NCW; ++j) {
    const int c = cw + WPB * j;
    if (c < NRB) {
#pragma unroll
      for (int v = 0; v < 4; ++v) Zh[(16 * rb + fk + v) * (L + 1) + 16 * c + fr] = cpx<float>{zhr[j][v], zhi[j][v]};
    }
  }
  __syncthreads();

  // --- P' = 2^-28 conj(A^') B^' for k1 <= L/2 (zero up to K1 - 1); this lane's entries
  // e = t + TPI i (k0 = e % L, k1 = e / L) are held until every wave has read Z^,
  // then staged column-major (PT[k1][k0])
  constexpr int NPE = L * K1 / TPI;
  static_assert(NPE * TPI == L * K1, "product entries per lane");
  float ppr[NPE], ppi[NPE];
#pragma unroll
  for (int i = 0; i < NPE; ++i) {
    const int e = t + TPI * i;
    const int k0 = e % L, k1 = e / L;
    float pr = 0.f, pi = 0.f;
    if (k1 <= L / 2) {
      const cpx<float> z = Zh[k0 * (L + 1) + k1];
      const cpx<float> n = Zh[((L - k0) % L) * (L + 1) + ((L - k1) % L)];
      // A^ = (z + conj n) / 2 and B^ = (z - conj n) / 2i, with the 1/4 in the scale
      const float ar = z.x + n.x, ai = z.y - n.y;
      const float br = z.y + n.y, bi = n.x - z.x;
      constexpr float s30 = 1.0f / 1073741824.0f;  // 2^-28 / 4
      pr = (ar * br + ai * bi) * s30;
      pi = (ar * bi - ai * br) * s30;
    }
    ppr[i] = pr;
    ppi[i] = pi;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NPE; ++i) {
    const int e = t + TPI * i;
    const int o = (e / L) * Ld + (e % L);
    mf_split1(ppr[i], S0[o], S0[P + o]);
    mf_split1(ppi[i], S0[2 * P + o], S0[3 * P + o]);
  }
  __syncthreads();

  // --- Y' = 2^-7 conj(W) P' for the rows s0 < ns0: Yr = Wr Pr + Wi Pi, Yi = Wr Pi - Wi Pr.
  // Held until every wave has read P, then staged row-major (Y[s0][k1], stride K1 + 8).
  constexpr int YLd = K1 + 8;  // 16-byte rows
  constexpr int NKT = K1 / 16;                  // kept-column tiles
  constexpr int NYW = (NKT + WPB - 1) / WPB;    // of them per wave
  const bool rows = 16 * rb < ns0;
  constexpr float ysc = 1.0f / (float)(1 << Cf::YS);
  mf_f4 yyr[NYW], yyi[NYW];
#pragma unroll
  for (int j = 0; j < NYW; ++j) yyr[j] = yyi[j] = zero4;
  if (rows) {
#pragma unroll
    for (int j = 0; j < NYW; ++j) {
      const int c = cw + WPB * j;
      if (c >= NKT) continue;
      mf_f4 yr = zero4, yr2 = zero4, yi = zero4, yi2 = zero4;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int kk = ks * 32 + fk8;
        const int ai = (16 * rb + fr) * L + kk;
        const mf_h8 wrh = mf_ld8(Wrh + ai), wrl = mf_ld8(Wrl + ai);
        const mf_h8 wih = mf_ld8(Wih + ai), wil = mf_ld8(Wil + ai);
        const mf_h8 wnh = mf_ld8(Wnh + ai), wnl = mf_ld8(Wnl + ai);
        const int bi = (16 * c + fr) * Ld + kk;
        const mf_h8 prh = mf_ld8(S0 + bi), prl = mf_ld8(S0 + P + bi);
        const mf_h8 pih = mf_ld8(S0 + 2 * P + bi), pil = mf_ld8(S0 + 3 * P + bi);
        mf_split8(wrh, wrl, prh, prl, yr, yr2);
        mf_split8(wih, wil, pih, pil, yr, yr2);
        mf_split8(wrh, wrl, pih, pil, yi, yi2);
        mf_split8(wnh, wnl, prh, prl, yi, yi2);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        yyr[j][v] = mf_sum(yr[v], yr2[v]) * ysc;
        yyi[j][v] = mf_sum(yi[v], yi2[v]) * ysc;
      }
    }
  }
  __syncthreads();
  if (rows) {
#pragma unroll
    for (int j = 0; j < NYW; ++j) {
      const int c = cw + WPB * j;
      if (c >= NKT) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int o = (16 * rb + fk + v) * YLd + 16 * c + fr;
        mf_split1(yyr[j][v], S0[o], S0[P + o]);
        mf_split1(yyi[j][v], S0[2 * P + o], S0[3 * P + o]);
      }
    }
  }
  __syncthreads();

  // --- C' = Yr' Cr + Yi' Ci over k1 < K1, for s0 < ns0 and s1 < ns1
  mf_f4 acc[NCW], acc2[NCW];
#pragma unroll
  for (int j = 0; j < NCW; ++j) acc[j] = acc2[j] = zero4;
  if (rows) {
#pragma unroll
    for (int ks = 0; ks < K1 / 32; ++ks) {  // 32-deep steps
      const int ai = (16 * rb + fr) * YLd + ks * 32 + fk8;
      const mf_h8 yrh = mf_ld8(S0 + ai), yrl = mf_ld8(S0 + P + ai);
      const mf_h8 yih = mf_ld8(S0 + 2 * P + ai), yil = mf_ld8(S0 + 3 * P + ai);
#pragma unroll
      for (int j = 0; j < NCW; ++j) {
        const int c = cw + WPB * j;
        if (c >= NRB || 16 * c >= ns1) continue;
        const int bi = (16 * c + fr) * K1 + ks * 32 + fk8;
        const mf_h8 crh = mf_ld8(tab.c[0] + bi), crl = mf_ld8(tab.c[1] + bi);
        const mf_h8 cih = mf_ld8(tab.c[2] + bi), cil = mf_ld8(tab.c[3] + bi);
        mf_split8(yrh, yrl, crh, crl, acc[j], acc2[j]);
        mf_split8(yih, yil, cih, cil, acc[j], acc2[j]);
      }
    }
    if constexpr (K1 % 32 != 0) {  // one 16-deep step
      constexpr int k16 = K1 - 16;
      const int ai = (16 * rb + fr) * YLd + k16 + fk;
      const mf_h4 yrh = mf_ld(S0 + ai), yrl = mf_ld(S0 + P + ai);
      const mf_h4 yih = mf_ld(S0 + 2 * P + ai), yil = mf_ld(S0 + 3 * P + ai);
#pragma unroll
      for (int j = 0; j < NCW; ++j) {
        const int c = cw + WPB * j;
        if (c >= NRB || 16 * c >= ns1) continue;
        const int bi = (16 * c + fr) * K1 + k16 + fk;
        const mf_h4 crh = mf_ld(tab.c[0] + bi), crl = mf_ld(tab.c[1] + bi);
        const mf_h4 cih = mf_ld(tab.c[2] + bi), cil = mf_ld(tab.c[3] + bi);
        mf_split(yrh, yrl, crh, crl, acc[j], acc2[j]);
        mf_split(yih, yil, cih, cil, acc[j], acc2[j]);
      }
    }
  }
  __syncthreads();  // every wave has read Y before the window overwrites it
  // C = C' 2^(2 XS + 35) / (sa sb L^2)
  const double out = ldexp(1.0, 2 * Cf::XS + 35) / ((double)sa * (double)sb * (double)(L * L));
  float* S = reinterpret_cast<float*>(region);
  if (rows) {
#pragma unroll
    for (int j = 0; j < NCW; ++j) {
      const int c = cw + WPB * j;
      if (c >= NRB) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int s0 = 16 * rb + fk + v, s1 = 16 * c + fr;
        if (s0 < ns0 && s1 < ns1) S[s1 * ns0 + s0] = (float)((double)mf_sum(acc[j][v], acc2[j][v]) * out);
      }
    }
  }
  __syncthreads();
}

}  // namespace mqpiv
