// mqpiv_wave2d.hpp -- split-precision matrix-core correlation of 2-D IAs, one IA
// per wave, every intermediate in registers.
//
// The same values as the FFT correlation of fftcc.jl:101-126 / fft_ops.jl:95-105
// (C(s) = sum_x F(x) G(x + s) over the fully-overlapping window s in [0, 2SM]),
// computed as four chained products with DFT matrices on the matrix cores
// (v_mfma_f32_16x16x32_f16).  With i the image's dim 1 and j its dim 2, both
// windows placed at the origin of an L x L buffer:
//
//   1. X(j, k)  = sum_i x(i, j) w1(i, k)       real -> complex along i at the
//      L/2 odd frequencies k + 1/2 (a negacyclic transform: wrapped terms change
//      sign, and none reaches the window, so the correlation is unchanged; the
//      L real inputs map to exactly L/2 complex outputs, with no DC or Nyquist
//      column to carry separately)
//   2. X^(q, k) = sum_j W(q, j) X(j, k)        complex along j, all L frequencies
//   3. P = conj(F^) G^                          element by element
//   4. Y(k, s1) = sum_q P(q, k) conj W(q, s1)   back along j, s1 < 16 NTS
//   5. C(s0, s1) = 2/L^2 Re sum_k Y(k, s1) exp(+i pi s0 (2k + 1) / L)
//
// A product's D tile leaves row index 4(l/16)+v in registers and column l%16 on
// lanes; the next product contracts over the register index, taking the tile
// as its B operand (stages 2, 5) or, transposed, as its A operand (stage 4).
// Two D tiles of consecutive rows give a 32-deep K fragment whose k order is
// a fixed permutation of the rows (wv_perm); the DFT fragments are stored in
// that order, so no tile is ever moved between lanes or through LDS.  A complex
// product takes four real products accumulated straight into its real and
// imaginary tiles (Re = Wr Xr + (-Wi) Xi, Im = Wi Xr + Wr Xi).
//
// Split operands: v = hi + lo with hi = fp16(v), lo = fp16(v - hi); a product is
// a_hi b_hi + a_hi b_lo + a_lo b_hi in one FP32 accumulator (a_lo b_lo, about
// 2^-22 relative, is dropped).  The windows and P are scaled by a power of two
// so that their largest magnitude lies in [2^4, 2^5) (a wave maximum per IA)
// and the DFT fragments by 4: every product's output (at most L 2^5 4 per
// real product pair) then lands in FP16's range as it is, with no rescaling
// between the stages, while the lo parts of all values above 2^-3 of the
// operands stay in FP16's normal range (smaller ones keep an absolute error
// of at most 2^-25, below 2^-27 of the operand's maximum).  The powers are
// undone exactly at the end.
//
// Fragment conventions of v_mfma_f32_16x16x32_f16 (lane l, g = l/16, c = l%16):
// A (16x32): A[c][8g + e]; B (32x16): B[8g + e][c]; D (16x16): D[4g + v][c].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mqpiv_kernels2d.hpp"

namespace mqpiv {

// Scheduling barriers between the products' fragment groups (1: on), which
// keep the compiler from hoisting every fragment load of a stage at once.
#ifndef MQPIV_WV_SB
#define MQPIV_WV_SB 1
#endif
__device__ __forceinline__ void wv_sb() {
#if MQPIV_WV_SB
  __builtin_amdgcn_sched_barrier(0);
#endif
}

// When ZNCC runs with a refine list, an IA goes to the FP64 transforms when its
// FP32-level values could move an output past the FP64 results' bar (integer
// peak identical, SN within 1e-4 relative, displacement within 0.01 px):
// * SN = max / avg, avg the mean over the SN window.  avg's error stayed below
//   3.6e-7 x mean|r| over ~12,000 IAs of four kinds of particle image (dense,
//   sparse, strong background and noise, a 100x illumination ramp; interSize
//   32 and 64), so |avg| >= kRefineSnK x mean|r| keeps SN's relative error
//   below 1.8e-5;
// * the peak: the two largest values of the peak window (or the largest and 0)
//   at least kRefinePeak apart, ten times the largest value error expected;
// * the Gaussian fit: |D| >= kRefineFit (1 + 4|q|) per dimension (D the fit's
//   denominator, q its offset), which bounds q's error by 1e-3 px for a log
//   error of 2e-5;
// * every value of the peak window finite, and every window energy there
//   keeping 24 bits of its non-negative terms.
#ifndef MQPIV_WV_BAND
#define MQPIV_WV_BAND 16
#endif
constexpr int kWvBand = MQPIV_WV_BAND;  // vector-field rows per band of the visiting order

constexpr double kRefineSnK = 0.02;
constexpr float kRefinePeak = 1e-4f;
constexpr double kRefineFit = 0.02;

typedef _Float16 wh8 __attribute__((ext_vector_type(8)));
typedef float wf4 __attribute__((ext_vector_type(4)));
typedef float wf8 __attribute__((ext_vector_type(8)));

// Row of a D-tile pair that k slot 8g + e of a 32-deep fragment holds.
__host__ __device__ constexpr int wv_perm(int g, int e) { return e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4); }

// Transform length L (a multiple of 32) and NTS output tiles (windows up to
// 16 NTS per dim).  FC: the IA size one kernel takes per dim.
template <int L, int NTS>
struct WvCfg {
  static_assert(L % 32 == 0 && L <= 128, "length a multiple of 32");
  static constexpr int NT = L / 16;             // 16-tiles along j and the frequencies q
  static constexpr int KS = L / 32;             // 32-deep steps over a length-L contraction
  static constexpr int NH = L / 32;             // tiles of the L/2 odd frequencies k
  static constexpr int FC = L >= 64 ? L - 32 : L;  // IA capacity per dim
  static constexpr int FT = FC / 16, FK = FC / 32;  // IA tiles along j, 32-steps along i
  static_assert(FT % 2 == 0, "IA tiles in pairs");
  static constexpr int nW1 = KS * NT * 2;
  static constexpr int nW2 = NT * KS * 4;
  static constexpr int nW4 = KS * NTS * 4;
  static constexpr int nW5 = NTS * KS * 2;
  static constexpr int o1 = 0, o2 = nW1, o4 = o2 + nW2, o5 = o4 + nW4, nF = o5 + nW5;
  static constexpr size_t table_bytes = (size_t)nF * 1024;  // a fragment: 64 lanes x 8 halfs

  static constexpr int LOG2L = L <= 32 ? 5 : (L <= 64 ? 6 : 7);  // ceil(log2 L)
  // fragment indices; p = 0 hi, 1 lo; q = 0 real, 1 imaginary (the negated
  // imaginary part is formed from it after the load: FP16 sign flips)
  static constexpr int f1(int ks, int tn, int p) { return o1 + (ks * NT + tn) * 2 + p; }
  static constexpr int f2(int tm, int ks, int q, int p) { return o2 + ((tm * KS + ks) * 2 + q) * 2 + p; }
  static constexpr int f4(int ks, int tns, int q, int p) { return o4 + ((ks * NTS + tns) * 2 + q) * 2 + p; }
  static constexpr int f5(int tms, int ks, int p) { return o5 + (tms * KS + ks) * 2 + p; }
};

// The DFT fragments, scaled by 4, as hi/lo FP16 pairs in lane order.
template <int L, int NTS>
struct alignas(16) WvTables {
  using C = WvCfg<L, NTS>;
  _Float16 f[C::nF][64][8];
  constexpr void put(int fr, int lane, int e, double v) {
    v *= 4.0;
    const _Float16 hi = (_Float16)v;
    f[fr][lane][e] = hi;
    f[fr + 1][lane][e] = (_Float16)(v - (double)hi);
  }
  constexpr WvTables() : f{} {
    constexpr Twiddles<L, double> tw{};       // cos, sin of 2 pi k / L
    constexpr Twiddles<2 * L, double> tw2{};  // cos, sin of pi k / L
    constexpr int H = L / 2;
    for (int lane = 0; lane < 64; ++lane) {
      const int g = lane / 16, c = lane % 16;
      for (int e = 0; e < 8; ++e) {
        // stage 1 B: row i = 32 ks + 8g + e, column slot 16 tn + c
        for (int ks = 0; ks < C::KS; ++ks)
          for (int tn = 0; tn < C::NT; ++tn) {
            const int i = 32 * ks + 8 * g + e, slot = 16 * tn + c;
            const int k = slot < H ? slot : slot - H;
            const int a = (i * (2 * k + 1)) % (2 * L);
            put(C::f1(ks, tn, 0), lane, e, slot < H ? tw2.c[a] : -tw2.s[a]);
          }
        // stage 2 A: row q = 16 tm + c, k slot j = 32 ks + perm(g, e); W = exp(-2 pi i q j / L)
        for (int tm = 0; tm < C::NT; ++tm)
          for (int ks = 0; ks < C::KS; ++ks) {
            const int q = 16 * tm + c, j = 32 * ks + wv_perm(g, e);
            const int a = (q * j) % L;
            put(C::f2(tm, ks, 0, 0), lane, e, tw.c[a]);
            put(C::f2(tm, ks, 1, 0), lane, e, -tw.s[a]);
          }
        // stage 4 B: k slot q = 32 ks + perm(g, e), column s1 = 16 tns + c; conj W = exp(+2 pi i q s1 / L)
        for (int ks = 0; ks < C::KS; ++ks)
          for (int tns = 0; tns < NTS; ++tns) {
            const int q = 32 * ks + wv_perm(g, e), s1 = 16 * tns + c;
            const int a = (q * s1) % L;
            put(C::f4(ks, tns, 0, 0), lane, e, tw.c[a]);
            put(C::f4(ks, tns, 1, 0), lane, e, tw.s[a]);
          }
        // stage 5 A: row s0 = 16 tms + c; k slot = (virtual tile 2 ks + e/4, row 4g + e%4):
        // virtual tiles [Re Y tiles 0..NH-1, Im Y tiles 0..NH-1]
        for (int tms = 0; tms < NTS; ++tms)
          for (int ks = 0; ks < C::KS; ++ks) {
            const int vt = 2 * ks + (e >= 4 ? 1 : 0);
            const int part = vt / C::NH, th = vt % C::NH;
            const int k = 16 * th + 4 * g + (e & 3), s0 = 16 * tms + c;
            const int a = (s0 * (2 * k + 1)) % (2 * L);
            put(C::f5(tms, ks, 0), lane, e, part == 0 ? tw2.c[a] : -tw2.s[a]);
          }
      }
    }
  }
};

static __device__ const WvTables<64, 3> kWvTab64{};
template <int L, int NTS>
__device__ __forceinline__ const uint4* wv_gtab() {
  static_assert(L == 64 && NTS == 3, "tables built for (64, 3)");
  return reinterpret_cast<const uint4*>(&kWvTab64);
}

__device__ __forceinline__ wf4 wv_mm(wh8 a, wh8 b, wf4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }

struct WvHl {
  wh8 h, l;
};
__device__ __forceinline__ WvHl wv_split(wf8 v) {
  WvHl r;
  r.h = __builtin_convertvector(v, wh8);
  r.l = __builtin_convertvector(v - __builtin_convertvector(r.h, wf8), wh8);
  return r;
}
__device__ __forceinline__ wf8 wv_cat(wf4 a, wf4 b) { return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7); }
__device__ __forceinline__ float wv_amax4(wf4 v) { return fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3]))); }
// wave maximum (every lane gets it)
__device__ __forceinline__ float wv_wave_max(float v) {
  Group<64>::steps([&](auto mv) { v = fmaxf(v, mv(0.0f, v)); });
  return bcast63(v);
}
// power of two k with m 2^k in [2^4, 2^5) (0 for m = 0), clamped to stay finite
__device__ __forceinline__ int wv_pow2(float m) {
  if (!(m > 0.0f)) return 0;
  int e = 0;
  (void)frexpf(m, &e);
  const int k = 5 - e;
  return k < -100 ? -100 : (k > 100 ? 100 : k);
}
// log out of line: its constants would otherwise be hoisted out of the IA loop
__device__ __noinline__ double wv_log(double x) { return log(x); }
__device__ __forceinline__ wh8 wv_frag(const uint4* T, int f, int lane) { return __builtin_bit_cast(wh8, T[f * 64 + lane]); }
// -x for a fragment of FP16 values (sign bits flipped)
__device__ __forceinline__ wh8 wv_neg(wh8 x) {
  uint4 u = __builtin_bit_cast(uint4, x);
  u.x ^= 0x80008000u;
  u.y ^= 0x80008000u;
  u.z ^= 0x80008000u;
  u.w ^= 0x80008000u;
  return __builtin_bit_cast(wh8, u);
}

// Per-wave LDS window bytes for ns0 x ns1 similarity values.
__host__ __device__ inline size_t wv_win_bytes(int ns0, int ns1) { return ((size_t)ns0 * ns1 * 4 + 15) & ~(size_t)15; }

// Waves per block (IAs in flight per CU) for IAs of NQ x NQ sub-blocks: one
// block keeps three waves per SIMD within 168 VGPRs; more blocks also hold the
// sum of their products (two waves per SIMD).
template <int NQ>
struct WvOcc {
  static constexpr int WAVES = NQ == 1 ? 12 : 8;
  static constexpr int OCC = WAVES / 4;
};

// One wave per IA, WAVES per block, a persistent grid: each XCD (hardware block
// b runs on XCD b % 8) takes a contiguous range of IAs, so overlapping search
// windows meet in its L2.  NQ: at most NQ x NQ sub-blocks of FC x FC per IA.
template <int L, int NTS, int ALG, int NQ>
__global__ __launch_bounds__((64 * WvOcc<NQ>::WAVES), (WvOcc<NQ>::OCC)) void piv2d_wave_kernel(
    Job j0, const Job* __restrict__ jobs, const int64_t* __restrict__ prefix, int n_jobs, int64_t n_ia_total,
    int ns0_cap, int ns1_cap, int32_t* __restrict__ refine) {
  using C = WvCfg<L, NTS>;
  constexpr int NT = C::NT, KS = C::KS, NH = C::NH, FT = C::FT, FK = C::FK;
  constexpr int WAVES = WvOcc<NQ>::WAVES;
  extern __shared__ __align__(16) unsigned char smem[];
  const uint4* T = reinterpret_cast<const uint4*>(smem);
  {
    const uint4* gtab = wv_gtab<L, NTS>();
    uint4* Tw = reinterpret_cast<uint4*>(smem);
    for (int i = (int)threadIdx.x; i < C::nF * 64; i += 64 * WAVES) Tw[i] = gtab[i];
  }
  __syncthreads();
  // the wave index, made wave-uniform for the compiler: every per-IA value below
  // is then a scalar
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  float* win = reinterpret_cast<float*>(smem + C::table_bytes + (size_t)w * wv_win_bytes(ns0_cap, ns1_cap));

  const int32_t nt = (int32_t)n_ia_total;
  const int32_t nb = (int32_t)gridDim.x, b = (int32_t)blockIdx.x;
  int32_t first, stride, end;
  if ((nb & 7) == 0) {
    const int32_t x = b & 7, per = nb >> 3;
    first = (int32_t)(((int64_t)x * nt) >> 3) + (b >> 3) * WAVES + w;
    end = (int32_t)(((int64_t)(x + 1) * nt) >> 3);
    stride = per * WAVES;
  } else {
    first = b * WAVES + w;
    end = nt;
    stride = nb * WAVES;
  }

#pragma unroll 1
  for (int32_t iag = first; iag < end; iag += stride) {
    // the lane index is re-read every iteration: values derived from it would
    // otherwise be hoisted out of the loop and held (spilled) across the body
    int lane = (int)threadIdx.x & 63;
    __asm__ volatile("" : "+v"(lane));
    const int g = lane >> 4, c = lane & 15;
    Job J = j0;
    int32_t ia32 = iag;
    if (jobs) {
      const int jb = find_job(prefix, n_jobs, iag);
      J = jobs[jb];
      ia32 = iag - (int32_t)prefix[jb];
    }
    const int sm0 = J.sm[0], sm1 = J.sm[1];
    const int is0 = J.is[0], is1 = J.is[1];
    const int ns0 = 2 * sm0 + 1, ns1 = 2 * sm1 + 1;
    const uint32_t vn0 = (uint32_t)J.vf_n[0];
    // visiting order: bands of kWvBand vector-field rows, each walked column by
    // column, so that the IAs in flight on an XCD cover a compact patch of the
    // image whose overlapping search regions its L2 holds
    int32_t v0, v1;
    {
      const uint32_t bw = (uint32_t)kWvBand * (uint32_t)J.vf_n[1];
      const uint32_t band = (uint32_t)ia32 / bw;
      const uint32_t r = (uint32_t)ia32 - band * bw;
      const uint32_t b0 = band * (uint32_t)kWvBand;
      const uint32_t h = vn0 - b0 < (uint32_t)kWvBand ? vn0 - b0 : (uint32_t)kWvBand;
      v1 = (int32_t)(r / h);
      v0 = (int32_t)(b0 + (r - (uint32_t)v1 * h));
    }
    const int64_t ia = (int64_t)v0 + (int64_t)vn0 * v1;  // the IA's index in the vector field
    const Geo g0 = geo_of(v0, is0, sm0, J.st[0], (int32_t)J.dims[0]);
    const Geo g1 = geo_of(v1, is1, sm1, J.st[1], (int32_t)J.dims[1]);
    const int dt = J.in_dtype;
    const int64_t d0 = J.dims[0];
    const uint32_t ud0 = (uint32_t)d0;

    // ---- the windows, one after the other: lane (g, c) holds, for j-tile tm
    // and i-step ks, the 8 values x(i = 32 ks + 8g + e, j = 16 tm + c).
    // Positions outside the IA / the clipped search region read an in-window
    // pixel and are zeroed (piv_ops.jl:99-127).  The search window runs its
    // first two products before the IA is loaded, so that one window's tiles
    // are held at a time (the larger first: fewer registers at the peak).
    const bool want_max = J.has_threshold != 0;
    const int64_t g_org = g0.sr + d0 * g1.sr, f_org = g0.tlf + d0 * g1.tlf;
    double sumF2 = 0.0, muG = 0.0;
    int skip = 0;

    // stage 1: X = x w1 (A = window rows j, B = table fragments)
    auto stage1 = [&](auto& xa, auto& d1, auto mtc, auto nkc) {
      constexpr int MT = decltype(mtc)::value, NK = decltype(nkc)::value;
      const wf4 z4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int bb = 0; bb < NT; ++bb) d1[a][bb] = z4;
#pragma unroll
      for (int ks = 0; ks < NK; ++ks)
#pragma unroll
        for (int tn = 0; tn < NT; ++tn) {
          wv_sb();
          const wh8 bh = wv_frag(T, C::f1(ks, tn, 0), lane), bl = wv_frag(T, C::f1(ks, tn, 1), lane);
#pragma unroll
          for (int tm = 0; tm < MT; ++tm) d1[tm][tn] = wv_mm(xa[tm][ks].h, bh, d1[tm][tn]);
#pragma unroll
          for (int tm = 0; tm < MT; ++tm) d1[tm][tn] = wv_mm(xa[tm][ks].h, bl, d1[tm][tn]);
#pragma unroll
          for (int tm = 0; tm < MT; ++tm) d1[tm][tn] = wv_mm(xa[tm][ks].l, bh, d1[tm][tn]);
        }
    };
    // stage 2: X^ = W X (A = table fragments, B = the stage-1 tiles), complex
    // products as four real ones; consume(tm, th, Re, Im) takes each finished tile
    auto stage2 = [&](auto& d1, auto nkc, auto&& consume) {
      constexpr int NK = decltype(nkc)::value;
      WvHl xr[NK][NH], xi[NK][NH];
#pragma unroll
      for (int ks = 0; ks < NK; ++ks)
#pragma unroll
        for (int th = 0; th < NH; ++th) {
          xr[ks][th] = wv_split(wv_cat(d1[2 * ks][th], d1[2 * ks + 1][th]));
          xi[ks][th] = wv_split(wv_cat(d1[2 * ks][th + NH], d1[2 * ks + 1][th + NH]));
        }
#pragma unroll
      for (int tm = 0; tm < NT; ++tm) {
        const wf4 z4 = {0.f, 0.f, 0.f, 0.f};
        wf4 re[NH], im[NH];
#pragma unroll
        for (int th = 0; th < NH; ++th) re[th] = im[th] = z4;
#pragma unroll
        for (int ks = 0; ks < NK; ++ks) {
          wv_sb();
          const wh8 rh = wv_frag(T, C::f2(tm, ks, 0, 0), lane), rl = wv_frag(T, C::f2(tm, ks, 0, 1), lane);
          const wh8 ih = wv_frag(T, C::f2(tm, ks, 1, 0), lane), il = wv_frag(T, C::f2(tm, ks, 1, 1), lane);
          const wh8 nh = wv_neg(ih), nl = wv_neg(il);
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(rh, xr[ks][th].h, re[th]);
            im[th] = wv_mm(ih, xr[ks][th].h, im[th]);
          }
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(nh, xi[ks][th].h, re[th]);
            im[th] = wv_mm(rh, xi[ks][th].h, im[th]);
          }
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(rh, xr[ks][th].l, re[th]);
            im[th] = wv_mm(ih, xr[ks][th].l, im[th]);
          }
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(nh, xi[ks][th].l, re[th]);
            im[th] = wv_mm(rh, xi[ks][th].l, im[th]);
          }
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(rl, xr[ks][th].h, re[th]);
            im[th] = wv_mm(il, xr[ks][th].h, im[th]);
          }
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(nl, xi[ks][th].h, re[th]);
            im[th] = wv_mm(rl, xi[ks][th].h, im[th]);
          }
        }
#pragma unroll
        for (int th = 0; th < NH; ++th) consume(tm, th, re[th], im[th]);
      }
    };

    // Window loads.  Sub-block (oa, ob) of the IA: rows [oa, oa + FC), columns
    // [ob, ob + FC); its part of the search buffer: rows [oa, oa + L), columns
    // [ob, ob + L).  The loaders leave the values in fv / gv and the positions
    // inside the data in (mi, mj); full: every position is inside (uniform).
    auto load_g = [&](int oa, int ob, float (&gv)[NT][KS][8], uint32_t (&mi)[KS], bool (&mj)[NT]) -> bool {
      uint32_t ri[KS][8], cj[NT];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        mi[ks] = 0;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int x = oa + 32 * ks + 8 * g + e - g0.goff;
          mi[ks] |= (x >= 0 && x < g0.gsz ? 1u : 0u) << e;
          ri[ks][e] = (uint32_t)min(max(x, 0), g0.gsz - 1);
        }
      }
#pragma unroll
      for (int tm = 0; tm < NT; ++tm) {
        const int x = ob + 16 * tm + c - g1.goff;
        mj[tm] = x >= 0 && x < g1.gsz;
        cj[tm] = ud0 * (uint32_t)min(max(x, 0), g1.gsz - 1);
      }
      auto fetch = [&](auto tin) {
        using TIN = decltype(tin);
        const gptr<char> gb = as_global<char>(J.img2) + g_org * (int64_t)sizeof(TIN);
#pragma unroll
        for (int tm = 0; tm < NT; ++tm)
#pragma unroll
          for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int e = 0; e < 8; ++e)
              gv[tm][ks][e] = (float)*(gptr<TIN>)(gb + (cj[tm] + ri[ks][e]) * (uint32_t)sizeof(TIN));
      };
      // FP32 windows whose rows start at the buffer's first row, on 16-byte
      // boundaries, in whole groups of 4: each lane loads its 8 rows of a
      // column as two 16-byte vectors (a group past the data reads the first
      // group; the mask zeroes it)
      const bool vec = dt == DT_F32 && g0.goff == 0 && (g0.gsz & 3) == 0 && (g0.sr & 3) == 0 && (d0 & 3) == 0 &&
                       (((uintptr_t)J.img2) & 15) == 0;
      if (vec) {
        const gptr<char> gb = as_global<char>(J.img2) + g_org * (int64_t)sizeof(float);
#pragma unroll
        for (int tm = 0; tm < NT; ++tm)
#pragma unroll
          for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const int r = oa + 32 * ks + 8 * g + 4 * h;
              const uint32_t rr = r + 4 <= g0.gsz ? (uint32_t)r : 0u;
              const wf4 v = *(gptr<wf4>)(gb + (cj[tm] + rr) * (uint32_t)sizeof(float));
#pragma unroll
              for (int q = 0; q < 4; ++q) gv[tm][ks][4 * h + q] = v[q];
            }
      } else if (dt == DT_F32) {
        fetch(float{});
      } else {
        fetch(double{});
      }
      return g0.goff == 0 && g1.goff == 0 && oa + L <= g0.gsz && ob + L <= g1.gsz;
    };
    auto load_f = [&](int oa, int ob, float (&fv)[FT][FK][8], uint32_t (&mi)[FK], bool (&mj)[FT]) -> bool {
      uint32_t ri[FK][8], cj[FT];
#pragma unroll
      for (int ks = 0; ks < FK; ++ks) {
        mi[ks] = 0;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int i = oa + 32 * ks + 8 * g + e;
          mi[ks] |= (i < is0 ? 1u : 0u) << e;
          ri[ks][e] = (uint32_t)min(i, is0 - 1);
        }
      }
#pragma unroll
      for (int tm = 0; tm < FT; ++tm) {
        const int j = ob + 16 * tm + c;
        mj[tm] = j < is1;
        cj[tm] = ud0 * (uint32_t)min(j, is1 - 1);
      }
      auto fetch = [&](auto tin) {
        using TIN = decltype(tin);
        const gptr<char> fb = as_global<char>(J.img1) + f_org * (int64_t)sizeof(TIN);
#pragma unroll
        for (int tm = 0; tm < FT; ++tm)
#pragma unroll
          for (int ks = 0; ks < FK; ++ks)
#pragma unroll
            for (int e = 0; e < 8; ++e)
              fv[tm][ks][e] = (float)*(gptr<TIN>)(fb + (cj[tm] + ri[ks][e]) * (uint32_t)sizeof(TIN));
      };
      // as the search window: 16-byte row groups of 4 where they line up
      const bool vec = dt == DT_F32 && (is0 & 3) == 0 && (g0.tlf & 3) == 0 && (d0 & 3) == 0 &&
                       (((uintptr_t)J.img1) & 15) == 0;
      if (vec) {
        const gptr<char> fb = as_global<char>(J.img1) + f_org * (int64_t)sizeof(float);
#pragma unroll
        for (int tm = 0; tm < FT; ++tm)
#pragma unroll
          for (int ks = 0; ks < FK; ++ks)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const int r = oa + 32 * ks + 8 * g + 4 * h;
              const uint32_t rr = r + 4 <= is0 ? (uint32_t)r : 0u;
              const wf4 v = *(gptr<wf4>)(fb + (cj[tm] + rr) * (uint32_t)sizeof(float));
#pragma unroll
              for (int q = 0; q < 4; ++q) fv[tm][ks][4 * h + q] = v[q];
            }
      } else if (dt == DT_F32) {
        fetch(float{});
      } else {
        fetch(double{});
      }
      return oa + C::FC <= is0 && ob + C::FC <= is1;
    };
    // sum, sum of squares, maximum and minimum of the IA's values in one sub-block
    auto f_stats = [&](const float (&fv)[FT][FK][8], const uint32_t (&mi)[FK], const bool (&mj)[FT], bool full,
                       double& sf, double& sff, double& mx, double& mn) {
#pragma unroll
      for (int tm = 0; tm < FT; ++tm)
#pragma unroll
        for (int ks = 0; ks < FK; ++ks)
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const bool in = full || (mj[tm] && ((mi[ks] >> e) & 1u));
            const double f = in ? (double)fv[tm][ks][e] : 0.0;
            sf += f;
            sff += f * f;
            if (in) {
              mx = fmax(mx, f);
              mn = fmin(mn, f);
            }
          }
    };
    // the IA's statistics: ZNCC mean (zncc_ops.jl:10, in the images' element
    // type, Q12) and sum((F - mean)^2) by expansion, NSQECC sum(F^2), and the
    // maximum for the skip test (piv_ops.jl:134-157)
    float meanF = 0.f;
    auto f_finish = [&](double sf, double sff, double mx1) {
      if constexpr (ALG == ALG_ZNCC) {
        const int64_t nF = (int64_t)is0 * is1;
        const float m32 = (float)sf / (float)nF;
        const double muF = dt == DT_F32 ? (double)m32 : sf / (double)nF;
        meanF = dt == DT_F32 ? m32 : (float)muF;
        const double v = (sff - 2.0 * muF * sf) + (double)nF * muF * muF;
        sumF2 = v > 0.0 ? v : 0.0;
      } else if constexpr (ALG == ALG_NSQECC) {
        sumF2 = sff;
      }
      if (J.skip) skip = J.skip[ia] != 0;
      else if (want_max && mx1 < J.threshold) skip = 1;
    };

    // G (zncc: G - mean of the clipped search region), scaled: stages 1 and 2;
    // returns its scale
    float meanG = 0.f;
    auto g_phase = [&](int oa, int ob, auto&& consume) -> int {
      float gv[NT][KS][8];
      uint32_t mi[KS];
      bool mj[NT];
      const bool full = load_g(oa, ob, gv, mi, mj);
      auto in = [&](int tm, int ks, int e) { return full || (mj[tm] && ((mi[ks] >> e) & 1u)); };
      if constexpr (ALG == ALG_ZNCC && NQ == 1) {
        double sg = 0.0;
#pragma unroll
        for (int tm = 0; tm < NT; ++tm)
#pragma unroll
          for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int e = 0; e < 8; ++e) sg += in(tm, ks, e) ? (double)gv[tm][ks][e] : 0.0;
        sg = Group<64>::reduce(sg, OpAdd{}, nullptr, lane);
        const int64_t nG = (int64_t)g0.gsz * g1.gsz;
        const float m32 = (float)sg / (float)nG;
        muG = dt == DT_F32 ? (double)m32 : sg / (double)nG;
        meanG = dt == DT_F32 ? m32 : (float)muG;
      }
      float am = 0.f;
#pragma unroll
      for (int tm = 0; tm < NT; ++tm)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            float v = gv[tm][ks][e];
            if constexpr (ALG == ALG_ZNCC) v = v - meanG;
            v = in(tm, ks, e) ? v : 0.f;
            gv[tm][ks][e] = v;
            am = fmaxf(am, fabsf(v));
          }
      const int kG = wv_pow2(wv_wave_max(am));
      const float sG = ldexpf(1.0f, kG);
      WvHl ga[NT][KS];
#pragma unroll
      for (int tm = 0; tm < NT; ++tm)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          wf8 v;
#pragma unroll
          for (int e = 0; e < 8; ++e) v[e] = gv[tm][ks][e] * sG;
          ga[tm][ks] = wv_split(v);
        }
      wf4 d1g[NT][NT];
      stage1(ga, d1g, std::integral_constant<int, NT>{}, std::integral_constant<int, KS>{});
      stage2(d1g, std::integral_constant<int, KS>{}, consume);
      return kG;
    };
    // F (zncc: F - mean), scaled by 2^kfix (or, kfix = INT32_MIN, by its own
    // maximum): stages 1 and 2; returns the scale.  With one sub-block the IA's
    // statistics come from this load.
    auto f_phase = [&](int oa, int ob, int kfix, auto&& consume) -> int {
      float fv[FT][FK][8];
      uint32_t mi[FK];
      bool mj[FT];
      const bool full = load_f(oa, ob, fv, mi, mj);
      auto in = [&](int tm, int ks, int e) { return full || (mj[tm] && ((mi[ks] >> e) & 1u)); };
      if constexpr (NQ == 1) {
        double sf = 0.0, sff = 0.0, mx1 = -INFINITY, mn1 = INFINITY;
        f_stats(fv, mi, mj, full, sf, sff, mx1, mn1);
        Group<64>::reduce_n<3>(sf, sff, mx1, OpAdd{}, OpAdd{}, OpMax{}, nullptr, lane);
        f_finish(sf, sff, mx1);
      }
      float am = 0.f;
#pragma unroll
      for (int tm = 0; tm < FT; ++tm)
#pragma unroll
        for (int ks = 0; ks < FK; ++ks)
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            float f = fv[tm][ks][e];
            if constexpr (ALG == ALG_ZNCC) f = f - meanF;
            f = in(tm, ks, e) ? f : 0.f;
            fv[tm][ks][e] = f;
            am = fmaxf(am, fabsf(f));
          }
      const int kF = kfix != INT32_MIN ? kfix : wv_pow2(wv_wave_max(am));
      const float sF = ldexpf(1.0f, kF);
      WvHl fa[FT][FK];
#pragma unroll
      for (int tm = 0; tm < FT; ++tm)
#pragma unroll
        for (int ks = 0; ks < FK; ++ks) {
          wf8 v;
#pragma unroll
          for (int e = 0; e < 8; ++e) v[e] = fv[tm][ks][e] * sF;
          fa[tm][ks] = wv_split(v);
        }
      wf4 d1f[FT][NT];
      stage1(fa, d1f, std::integral_constant<int, FT>{}, std::integral_constant<int, FK>{});
      stage2(d1f, std::integral_constant<int, FT / 2>{}, consume);
      return kF;
    };

    wf4 pr[NT][NH], pi[NT][NH];  // P (one sub-block: G^ first)
    int kFG;                     // P's scale: 2^kFG conj(F^) G^
    if constexpr (NQ == 1) {
      // the search window first (the larger: fewer registers at the peak)
      const int kG = g_phase(0, 0, [&](int tm, int th, wf4 re, wf4 im) {
        pr[tm][th] = re;
        pi[tm][th] = im;
      });
      // P = conj(F^) G^ as each F^ tile completes
      const int kF = f_phase(0, 0, INT32_MIN, [&](int tm, int th, wf4 fr, wf4 fi) {
        const wf4 gr = pr[tm][th], gi = pi[tm][th];
        pr[tm][th] = fr * gr + fi * gi;
        pi[tm][th] = fr * gi - fi * gr;
      });
      kFG = kF + kG;
    } else {
      // sub-blocks of at most FC x FC: C(s) = sum over the blocks of their own
      // correlation with their part of the search window, so the products are
      // summed (P) and transformed back once.  The IA's statistics and the ZNCC
      // means come first: one pass over the IA, and the search region's sum
      // from the integral array of image 2.
      const int nq0 = (is0 + C::FC - 1) / C::FC, nq1 = (is1 + C::FC - 1) / C::FC;
      int kF;
      {
        double sf = 0.0, sff = 0.0, mx1 = -INFINITY, mn1 = INFINITY;
        for (int qb = 0; qb < nq1; ++qb)
          for (int qa = 0; qa < nq0; ++qa) {
            float fv[FT][FK][8];
            uint32_t mi[FK];
            bool mj[FT];
            const bool full = load_f(C::FC * qa, C::FC * qb, fv, mi, mj);
            f_stats(fv, mi, mj, full, sf, sff, mx1, mn1);
          }
        Group<64>::reduce_n<3>(sf, sff, mx1, OpAdd{}, OpAdd{}, OpMax{}, nullptr, lane);
        mn1 = -Group<64>::reduce(-mn1, OpMax{}, nullptr, lane);
        f_finish(sf, sff, mx1);
        // one scale for every block of the IA, from its extremes (after the
        // mean, for ZNCC)
        const double am = fmax(fabs(mx1 - (double)meanF), fabs(mn1 - (double)meanF));
        kF = wv_pow2((float)am);
      }
      if constexpr (ALG == ALG_ZNCC) {
        const int64_t D0 = d0 + 1;
        const double* S1 = J.sat1;
        const int64_t a0 = g0.sr, b0 = g0.sr + g0.gsz, a1 = g1.sr, b1 = g1.sr + g1.gsz;
        const double sg = ((S1[b0 + D0 * b1] - S1[a0 + D0 * b1]) - S1[b0 + D0 * a1]) + S1[a0 + D0 * a1];
        const int64_t nG = (int64_t)g0.gsz * g1.gsz;
        const float m32 = (float)sg / (float)nG;
        muG = dt == DT_F32 ? (double)m32 : sg / (double)nG;
        meanG = dt == DT_F32 ? m32 : (float)muG;
      }
      const wf4 z4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int tm = 0; tm < NT; ++tm)
#pragma unroll
        for (int th = 0; th < NH; ++th) pr[tm][th] = pi[tm][th] = z4;
      int kG0 = 0;
      for (int qb = 0; qb < nq1; ++qb)
        for (int qa = 0; qa < nq0; ++qa) {
          wf4 gq[NT][NH], hq[NT][NH];  // this block's G^
          const int kG = g_phase(C::FC * qa, C::FC * qb, [&](int tm, int th, wf4 re, wf4 im) {
            gq[tm][th] = re;
            hq[tm][th] = im;
          });
          // P keeps the first block's scale; later blocks are brought to it (a
          // power of two, exact)
          if (qa == 0 && qb == 0) kG0 = kG;
          const float rs = ldexpf(1.0f, kG0 - kG);
          f_phase(C::FC * qa, C::FC * qb, kF, [&](int tm, int th, wf4 fr, wf4 fi) {
            const wf4 gr = gq[tm][th], gi = hq[tm][th];
            pr[tm][th] += (fr * gr + fi * gi) * rs;
            pi[tm][th] += (fr * gi - fi * gr) * rs;
          });
        }
      kFG = kF + kG0;
    }

    // ---- stage 4: Y = P^T conj W (A = P, transposed by the tile layout; B = table)
    float amP = 0.f;
#pragma unroll
    for (int tm = 0; tm < NT; ++tm)
#pragma unroll
      for (int th = 0; th < NH; ++th) amP = fmaxf(amP, fmaxf(wv_amax4(pr[tm][th]), wv_amax4(pi[tm][th])));
    const int kP = wv_pow2(wv_wave_max(amP));
    const float sP = ldexpf(1.0f, kP);
    wf4 yr[NH][NTS], yi[NH][NTS];
    {
      WvHl ar[NH][KS], ai[NH][KS];
#pragma unroll
      for (int th = 0; th < NH; ++th)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          ar[th][ks] = wv_split(wv_cat(pr[2 * ks][th], pr[2 * ks + 1][th]) * sP);
          ai[th][ks] = wv_split(wv_cat(pi[2 * ks][th], pi[2 * ks + 1][th]) * sP);
        }
#pragma unroll
      for (int tns = 0; tns < NTS; ++tns) {
        const wf4 z4 = {0.f, 0.f, 0.f, 0.f};
        wf4 re[NH], im[NH];
#pragma unroll
        for (int th = 0; th < NH; ++th) re[th] = im[th] = z4;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          wv_sb();
          const wh8 rh = wv_frag(T, C::f4(ks, tns, 0, 0), lane), rl = wv_frag(T, C::f4(ks, tns, 0, 1), lane);
          const wh8 ih = wv_frag(T, C::f4(ks, tns, 1, 0), lane), il = wv_frag(T, C::f4(ks, tns, 1, 1), lane);
          const wh8 nh = wv_neg(ih), nl = wv_neg(il);
          // Re Y = Pr Cr + Pi (-Ci), Im Y = Pr Ci + Pi Cr
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(ar[th][ks].h, rh, re[th]);
            im[th] = wv_mm(ar[th][ks].h, ih, im[th]);
          }
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(ai[th][ks].h, nh, re[th]);
            im[th] = wv_mm(ai[th][ks].h, rh, im[th]);
          }
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(ar[th][ks].h, rl, re[th]);
            im[th] = wv_mm(ar[th][ks].h, il, im[th]);
          }
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(ai[th][ks].h, nl, re[th]);
            im[th] = wv_mm(ai[th][ks].h, rl, im[th]);
          }
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(ar[th][ks].l, rh, re[th]);
            im[th] = wv_mm(ar[th][ks].l, ih, im[th]);
          }
#pragma unroll
          for (int th = 0; th < NH; ++th) {
            re[th] = wv_mm(ai[th][ks].l, nh, re[th]);
            im[th] = wv_mm(ai[th][ks].l, rh, im[th]);
          }
        }
#pragma unroll
        for (int th = 0; th < NH; ++th) {
          yr[th][tns] = re[th];
          yi[th][tns] = im[th];
        }
      }
    }

    // ---- stage 5: C = W5 Y (A = table, B = Y over the virtual tiles [Re, Im])
    wf4 cc[NTS][NTS];
    {
      auto vtile = [&](int vt, int tns) -> wf4 { return vt < NH ? yr[vt][tns] : yi[vt - NH][tns]; };
      WvHl yo[KS][NTS];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int tns = 0; tns < NTS; ++tns) yo[ks][tns] = wv_split(wv_cat(vtile(2 * ks, tns), vtile(2 * ks + 1, tns)));
      const wf4 z4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int a = 0; a < NTS; ++a)
#pragma unroll
        for (int bb = 0; bb < NTS; ++bb) cc[a][bb] = z4;
#pragma unroll
      for (int tms = 0; tms < NTS; ++tms)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          wv_sb();
          const wh8 ah = wv_frag(T, C::f5(tms, ks, 0), lane), al = wv_frag(T, C::f5(tms, ks, 1), lane);
#pragma unroll
          for (int tns = 0; tns < NTS; ++tns) cc[tms][tns] = wv_mm(ah, yo[ks][tns].h, cc[tms][tns]);
#pragma unroll
          for (int tns = 0; tns < NTS; ++tns) cc[tms][tns] = wv_mm(ah, yo[ks][tns].l, cc[tms][tns]);
#pragma unroll
          for (int tns = 0; tns < NTS; ++tns) cc[tms][tns] = wv_mm(al, yo[ks][tns].h, cc[tms][tns]);
        }
    }
    // C = cc 2 / L^2 / (2^(kFG + kP) 4^6): the operand scales, and the 4 of
    // each of the six table products in the chain
    const double cscale = ldexp(2.0 / ((double)L * (double)L), -(kFG + kP + 12));
#pragma unroll
    for (int tms = 0; tms < NTS; ++tms)
#pragma unroll
      for (int tns = 0; tns < NTS; ++tns)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int s0 = 16 * tms + 4 * g + v, s1 = 16 * tns + c;
          if (s0 < ns0 && s1 < ns1) win[s0 + ns0 * s1] = cc[tms][tns][v];
        }
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");

    // ---- normalisation, peak and SN (as the fused kernel: common_ops.jl:98-124,
    // piv_ops.jl:336-367, nsqecc_ops.jl:58-73, zncc_ops.jl:39-48)
    const int lo0 = g0.goff, hi0 = 2 * sm0 - g0.off_hi;
    const int lo1 = g1.goff, hi1 = 2 * sm1 - g1.off_hi;
    const int shi0 = 2 * sm0 - 1 - g0.off_hi, shi1 = 2 * sm1 - 1 - g1.off_hi;
    const bool want_sn = J.compute_sn != 0;
    const int nE = ns0 * ns1;
    Best<float> bst{0.f, INT32_MAX};
    double mx = -INFINITY, smv = 0.0;
    // ZNCC with a refine list: what decides whether this IA's FP32-level values
    // meet the FP64 results' bar -- the two largest values of the peak window,
    // the sum of |values| over the SN window, and whether a value there is not
    // finite or a window energy lost its digits to cancellation
    const bool rf = ALG == ALG_ZNCC && refine != nullptr;
    float t1 = -INFINITY, t2 = -INFINITY;
    double sab = 0.0;
    bool rbad = false;
    const int st0 = 64 % ns0, st1 = 64 / ns0;
    int w0 = lane % ns0, w1 = lane / ns0;
    auto step = [&]() {
      w0 += st0;
      w1 += st1;
      if (w0 >= ns0) {
        w0 -= ns0;
        ++w1;
      }
    };
    auto visit = [&](int e, int p0, int p1, float v) {
      const bool in_pk = p0 >= lo0 && p0 <= hi0 && p1 >= lo1 && p1 <= hi1;
      if (in_pk && v > bst.v) bst = Best<float>{v, e};
      if constexpr (ALG == ALG_ZNCC) {
        if (rf && in_pk) {
          rbad |= !(__builtin_fabsf(v) <= 3.4028235e38f);
          t2 = fmaxf(t2, fminf(t1, v));
          t1 = fmaxf(t1, v);
        }
      }
      if (want_sn && p0 >= lo0 && p0 <= shi0 && p1 >= lo1 && p1 <= shi1) {
        const double dv = (double)v;
        mx = (dv > mx || dv != dv) ? dv : mx;
        smv += dv;
        if constexpr (ALG == ALG_ZNCC) {
          if (rf) sab += __builtin_fabs(dv);
        }
      }
    };
    if constexpr (ALG == ALG_ZNCC || ALG == ALG_NSQECC) {
      const int64_t D0 = d0 + 1;
      const float sqFf = (float)sqrt(sumF2);
      const double mu = muG;
      const int64_t corner0 = (g0.sr + D0 * g1.sr) * (int64_t)sizeof(double);
      const gptr<char> sat2 = as_global<char>(J.sat2) + corner0;
      const gptr<char> sat1 = as_global<char>(J.sat1) + corner0;
      const uint32_t rowb = (uint32_t)D0 * (uint32_t)sizeof(double);
      auto at = [](gptr<char> bp, uint32_t off) { return *(gptr<double>)(bp + off); };
      constexpr int NB = 4;
      constexpr int NQC = (ALG == ALG_ZNCC) ? 2 : 1;
      for (int e0 = lane; e0 < nE; e0 += NB * 64) {
        double q[NB][NQC][4];
        double nbox[NB];
        int pw0[NB], pw1[NB];
#pragma unroll
        for (int jj = 0; jj < NB; ++jj) {
          const int e = e0 + jj * 64;
          pw0[jj] = w0;
          pw1[jj] = w1;
          step();
          const int s0 = e < nE ? pw0[jj] : 0, s1 = e < nE ? pw1[jj] : 0;
          const int lo0w = s0 > g0.goff ? s0 : g0.goff;
          int hi0w = (s0 + is0) < (g0.goff + g0.gsz) ? (s0 + is0) : (g0.goff + g0.gsz);
          hi0w = hi0w > lo0w ? hi0w : lo0w;
          const int lo1w = s1 > g1.goff ? s1 : g1.goff;
          int hi1w = (s1 + is1) < (g1.goff + g1.gsz) ? (s1 + is1) : (g1.goff + g1.gsz);
          hi1w = hi1w > lo1w ? hi1w : lo1w;
          const uint32_t a0 = (uint32_t)(lo0w - g0.goff) * (uint32_t)sizeof(double);
          const uint32_t b0 = (uint32_t)(hi0w - g0.goff) * (uint32_t)sizeof(double);
          const uint32_t a1 = (uint32_t)(lo1w - g1.goff) * rowb, b1 = (uint32_t)(hi1w - g1.goff) * rowb;
          const uint32_t c11 = b0 + b1, c01 = a0 + b1, c10 = b0 + a1, c00 = a0 + a1;
          q[jj][0][0] = at(sat2, c11);
          q[jj][0][1] = at(sat2, c01);
          q[jj][0][2] = at(sat2, c10);
          q[jj][0][3] = at(sat2, c00);
          if constexpr (ALG == ALG_ZNCC) {
            q[jj][NQC - 1][0] = at(sat1, c11);
            q[jj][NQC - 1][1] = at(sat1, c01);
            q[jj][NQC - 1][2] = at(sat1, c10);
            q[jj][NQC - 1][3] = at(sat1, c00);
          }
          nbox[jj] = (double)((hi0w - lo0w) * (hi1w - lo1w));
        }
#pragma unroll
        for (int jj = 0; jj < NB; ++jj) {
          const int e = e0 + jj * 64;
          if (e >= nE) break;
          double W = ((q[jj][0][0] - q[jj][0][1]) - q[jj][0][2]) + q[jj][0][3];
          if constexpr (ALG == ALG_ZNCC) {
            const double S1 = ((q[jj][NQC - 1][0] - q[jj][NQC - 1][1]) - q[jj][NQC - 1][2]) + q[jj][NQC - 1][3];
            const double G2 = W, M2 = nbox[jj] * mu * mu;
            W = (W - 2.0 * mu * S1) + M2;
            // the energy keeps at least 24 bits of its non-negative terms
            if (rf && pw0[jj] >= lo0 && pw0[jj] <= hi0 && pw1[jj] >= lo1 && pw1[jj] <= hi1)
              rbad |= !(W * 0x1p24 >= G2 + M2);
          }
          W = W > 0.0 ? W : 0.0;
          const double cv = (double)win[e] * cscale;
          // the cancelling sums in FP64; the ratio, stored in FP32, from FP32
          // operands (hardware square root and reciprocal, within 2 ulp)
          const float den = __builtin_amdgcn_sqrtf((float)W) * sqFf;
          float r;
          if constexpr (ALG == ALG_ZNCC) {
            r = (float)cv * __builtin_amdgcn_rcpf(den);  // zncc_ops.jl:39-48
          } else {
            const double num = (sumF2 - 2.0 * cv) + W;           // nsqecc_ops.jl:58-60
            r = den * __builtin_amdgcn_rcpf((float)num + den);  // nsqecc_ops.jl:72-73
          }
          win[e] = r;
          visit(e, pw0[jj], pw1[jj], r);
        }
      }
    } else {
      for (int e = lane; e < nE; e += 64) {
        const float r = (float)((double)win[e] * cscale);
        win[e] = r;
        visit(e, w0, w1, r);
        step();
      }
    }
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    Group<64>::steps([&](auto mv) {
      const Best<float> o{mv(0.0f, bst.v), mv((int32_t)INT32_MAX, bst.i)};
      bst = best_of(bst, o);
      mx = OpMaxNaN{}(mx, mv(OpMaxNaN::id(), mx));
      smv += mv(0.0, smv);
    });
    bst.v = bcast63(bst.v);
    bst.i = bcast63(bst.i);
    mx = bcast63(mx);
    smv = bcast63(smv);
    if constexpr (ALG == ALG_ZNCC) {
      if (rf) {
        Group<64>::steps([&](auto mv) {
          const float o1 = mv(-INFINITY, t1), o2 = mv(-INFINITY, t2);
          t2 = fmaxf(fminf(t1, o1), fmaxf(t2, o2));
          t1 = fmaxf(t1, o1);
          sab += mv(0.0, sab);
        });
        t1 = bcast63(t1);
        t2 = bcast63(t2);
        sab = bcast63(sab);
        rbad = __ballot(rbad) != 0;
      }
    }
    double sn = 0.0;
    const int u0 = shi0 - lo0 + 1, u1 = shi1 - lo1 + 1;
    const int nsn = (u0 > 0 && u1 > 0) ? u0 * u1 : 0;
    if (want_sn) sn = nsn > 0 ? mx / ((smv - mx) / (double)(nsn - 1)) : (double)NAN;

    // ---- Gaussian subpixel + write (piv_ops.jl:249-326, 236-243): lanes 0-8 hold
    // the 3x3 neighbourhood, as in the fused kernel
    const bool has_peak = !skip && bst.i != INT32_MAX;
    const int pk0 = has_peak ? bst.i % ns0 : 0, pk1 = has_peak ? bst.i / ns0 : 0;
    const bool interior = has_peak && (pk0 > lo0) && (pk0 < hi0) && (pk1 > lo1) && (pk1 < hi1);
    double lfit[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (interior) {
      const int tt = lane < 9 ? lane : 4;
      const double v = (double)win[(pk0 + tt % 3 - 1) + ns0 * (pk1 + tt / 3 - 1)];
      double mn = v;
#pragma unroll
      for (int m = 1; m < 16; m <<= 1) {
        const double o = __shfl_xor(mn, m);
        mn = (mn != mn || o != o) ? (double)NAN : (o < mn ? o : mn);
      }
      const double lv = wv_log(1.0 + v - mn);
      lfit[0] = __shfl(lv, 4);
      lfit[1] = __shfl(lv, 3);
      lfit[2] = __shfl(lv, 5);
      lfit[3] = __shfl(lv, 1);
      lfit[4] = __shfl(lv, 7);
    }
    if (lane == 0) {
      double r0, r1;
      if (skip) {
        r0 = r1 = 0.0;
        sn = 0.0;
      } else if (!has_peak) {
        r0 = -(double)(sm0 + 1);
        r1 = -(double)(sm1 + 1);
      } else {
        r0 = (double)(pk0 + 1) - (double)(sm0 + 1);
        r1 = (double)(pk1 + 1) - (double)(sm1 + 1);
        if (interior) {
          const double lmx = lfit[0], up = lfit[1], dn = lfit[2], lf = lfit[3], rt = lfit[4];
          const double D0 = 2 * up - 4 * lmx + 2 * dn, D1 = 2 * lf - 4 * lmx + 2 * rt;
          const double q0 = (up - dn) / D0;
          const double q1 = (lf - rt) / D1;
          r0 = ((double)(pk0 + 1) + q0) - (double)(sm0 + 1);
          r1 = ((double)(pk1 + 1) + q1) - (double)(sm1 + 1);
          if constexpr (ALG == ALG_ZNCC) {
            // an error e in the logs moves q by at most e (2 + 8 |q|) / |D|
            if (rf)
              rbad |= !(__builtin_fabs(D0) >= kRefineFit * (1.0 + 4.0 * __builtin_fabs(q0))) ||
                      !(__builtin_fabs(D1) >= kRefineFit * (1.0 + 4.0 * __builtin_fabs(q1)));
          }
        }
      }
      if constexpr (ALG == ALG_ZNCC) {
        if (rf && !skip) {
          bool f = rbad;
          f |= !(__builtin_fabsf(t1) >= kRefinePeak) || !((double)t1 - (double)t2 >= kRefinePeak);
          if (want_sn && nsn > 1) f |= !(__builtin_fabs((smv - mx) / (double)(nsn - 1)) >= kRefineSnK * sab / nsn);
          if (f) {
            const int32_t pos = atomicAdd(refine, 1);
            refine[kRefineHead + pos] = iag - ia32 + (int32_t)ia;
          }
        }
      }
      if (J.peaks) {
        J.peaks[2 * ia + 0] = has_peak ? pk0 + 1 : 0;
        J.peaks[2 * ia + 1] = has_peak ? pk1 + 1 : 0;
      }
      if (J.out_dtype == DT_F32) {
        float* vf = (float*)J.vf;
        vf[2 * ia + 0] = (float)r0;
        vf[2 * ia + 1] = (float)r1;
        if (J.sn) ((float*)J.sn)[ia] = (float)sn;
      } else {
        double* vf = (double*)J.vf;
        vf[2 * ia + 0] = r0;
        vf[2 * ia + 1] = r1;
        if (J.sn) ((double*)J.sn)[ia] = sn;
      }
    }
    // the window is rewritten by the next IA: every lane's reads are done
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  }
}

}  // namespace mqpiv
