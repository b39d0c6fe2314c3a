// mqpiv_kernels3d.hpp -- 3-D cross-correlation PIV on gfx950.
//
// A 3-D window (L^3 complex, 884 KB at L = 48 in FP32) does not fit in LDS, so
// each chunk of IAs runs four kernels through a device scratch buffer sized to
// stay in the last-level cache:
//   K1 slab_fwd  one block per (IA, z-slab i2): pack Z = A + iB for the slab,
//                forward 2-D transform (dims 1, 2) in registers + LDS, store
//                the slab; per-slab partial sum(F^2) and IA maxima; for ZNCC
//                and NSQECC the slab's 2-D window sums of G^2 (every window
//                position) from a summed-area table of the slab in LDS.
//   K2 zline     one block per (IA, pair of planes k1 / -k1): transform along
//                dim 3, P = conj(A^) B^ with the partner bin from the paired
//                plane, inverse along dim 3 keeping s2 < ns2.
//   K3 slab_inv  one block per (IA, s2): inverse 2-D transform down to the
//                (ns0 x ns1) window of real correlation values.
//   K4 finalize  one block per IA: FP64 window sums along dim 3 from the
//                slabs' window sums, ZNCC/NSQECC normalisation, argmax, 3-D
//                Gaussian subpixel, SN, VF/SN in the reference layout.
//
// Reference behaviour restated (files relative to the reference repo root):
//   coordinates/offsets  src/utils/piv_ops.jl:38-85 (incl. SR_TLF_off, :72)
//   correlation          src/crosscorrelation_algorithms/fftcc.jl:101-126
//   ZNCC / NSQECC        zncc_ops.jl:3-51, nsqecc_ops.jl:1-26
//   3-D window sums      src/utils/common_ops.jl:164-180, 223-245
//   peak                 src/utils/common_ops.jl:98-124 (z, c, r loop order)
//   subpixel             src/utils/piv_ops.jl:282-326 (gaussian_3D)
//   SN                   src/utils/piv_ops.jl:333-367; the reference's 3-D form
//                        throws (4 index ranges on a 3-D array), so 3-D SN uses
//                        the natural extension [1+o-, 2SM-o+] on every dim.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "mqpiv_kernels2d.hpp"

namespace mqpiv {

template <int L, typename T, int LPL>
struct Cfg3 {
  static constexpr int M = L / LPL;
  static constexpr int LD = L + 1;
  static constexpr int TPL = ((L * LPL + 63) / 64) * 64;  // threads for one plane of lines
  static constexpr int NT2 = ((2 * L * LPL + 63) / 64) * 64;  // K2: two planes
};

// Per-chunk scratch layout (all per IA of the chunk):
//   W    : L^3 complex T, [i2|k2|s2][k1][k0]
//   S    : ns0*ns1*ns2 T  (correlation window, column-major)
//   S2   : ns0*ns1*ns2 T  (masked NSQECC: the second correlation, mask x G^2)
//   part : L doubles x 3 (sum F^2, max img1, max mask) per slab
//   box  : L x ns0*ns1 doubles, [i2][s0 + ns0 s1]: slab i2's sum of G^2 over the
//          window at (s0, s1) (ZNCC: of G'^2, G' = G - mean) -- ZNCC and NSQECC
struct Scratch3 {
  void* w;
  void* s;
  void* s2;  // masked NSQECC only
  double* part;
  double* box;
  int64_t w_stride;  // elements of cpx<T> per IA
  int64_t s_stride;  // elements of T per IA
  int64_t box_stride;  // doubles per IA
};

struct Geo3 {
  Geo g[3];
  int32_t d[3];  // search-region shifts (multipass; 0 without)
};

__device__ __forceinline__ Geo3 geo3_of(const Job& J, int64_t ia) {
  Geo3 r;
  // 32-bit (the host keeps IA counts below 2^31)
  const uint32_t i = (uint32_t)ia, n0 = (uint32_t)J.vf_n[0], n1 = (uint32_t)J.vf_n[1];
  const uint32_t rest = i / n0;
  const uint32_t v0 = i - rest * n0;
  const uint32_t v2 = rest / n1;
  const uint32_t v1 = rest - v2 * n1;
  const int32_t v[3] = {(int32_t)v0, (int32_t)v1, (int32_t)v2};
  for (int k = 0; k < 3; ++k) {
    r.d[k] = shift_of(J, (int32_t)i, k, v[k]);
    r.g[k] = geo_of(v[k], J.is[k], J.sm[k], J.st[k], (int32_t)J.dims[k], r.d[k]);
  }
  return r;
}

// Column-major position (a, b, c) in an n0 x n1 x (any) box of the elements
// e = first + m * stride, m = 0, 1, ...: the divisions once, then a step with
// carries.
struct BoxWalk {
  int a, b, c, sa, sb, sc, n0, n1;
  __device__ __forceinline__ BoxWalk(int first, int stride, int n0_, int n1_)
      : n0(n0_ > 0 ? n0_ : 1), n1(n1_ > 0 ? n1_ : 1) {
    a = first % n0;
    const int r = first / n0;
    b = r % n1;
    c = r / n1;
    sa = stride % n0;
    const int q = stride / n0;
    sb = q % n1;
    sc = q / n1;
  }
  __device__ __forceinline__ void step() {
    a += sa;
    b += sb;
    c += sc;
    if (a >= n0) {
      a -= n0;
      ++b;
    }
    if (b >= n1) {
      b -= n1;
      ++c;
    }
  }
};

// Box sum of a 3-D FP64 integral array over [a, b) per dim.
__device__ __forceinline__ double sat_box3(const double* __restrict__ S, int64_t D0, int64_t D01, int64_t a0,
                                           int64_t b0, int64_t a1, int64_t b1, int64_t a2, int64_t b2) {
  const gptr<double> g = as_global<double>(S);
  auto at = [&](int64_t i, int64_t j, int64_t k) { return g[i + D0 * j + D01 * k]; };
  return (((at(b0, b1, b2) - at(a0, b1, b2)) - at(b0, a1, b2)) - at(b0, b1, a2)) +
         ((at(a0, a1, b2) + at(a0, b1, a2)) + at(b0, a1, a2)) - at(a0, a1, a2);
}

// ZNCC means of one IA (sum over the IA and over the clipped search region),
// rounded in the input element type like the reference (zncc_ops.jl:10, 22).
// The sums come from the per-IA sums (J.sat1 NULL; s1 = sums, image 1 at [ia],
// image 2 at [n_ia + ia]) or, for shifted search regions, from the integral
// arrays of both images (s1 = image 1's, J.sat1 = image 2's).
struct Means3 {
  float f32, g32;
  double f64, g64;
};
__device__ __forceinline__ Means3 means3(const Job& J, const Geo3& G, const double* s1, int64_t ia) {
  const int64_t D0 = J.dims[0] + 1, D01 = D0 * (J.dims[1] + 1);
  const Geo *g0 = &G.g[0], *g1 = &G.g[1], *g2 = &G.g[2];
  double sf, sg;
  if (!J.sat1) {
    sf = s1[ia];
    sg = s1[J.n_ia + ia];
  } else {
    sf = sat_box3(s1, D0, D01, g0->tlf, g0->tlf + J.is[0], g1->tlf, g1->tlf + J.is[1], g2->tlf, g2->tlf + J.is[2]);
    sg = sat_box3(J.sat1, D0, D01, g0->sr, g0->sr + g0->gsz, g1->sr, g1->sr + g1->gsz, g2->sr, g2->sr + g2->gsz);
  }
  const double nF = (double)J.is[0] * J.is[1] * J.is[2];
  const double nG = (double)g0->gsz * g1->gsz * g2->gsz;
  Means3 m;
  m.f32 = (float)sf / (float)nF;
  m.g32 = (float)sg / (float)nG;
  m.f64 = sf / nF;
  m.g64 = sg / nG;
  return m;
}

// ---------------------------------------------------------------------------
// K1: forward 2-D transform of slab i2 of the packed window.  Masked NSQECC
// (mask_nsqecc.jl:141-160) runs the pipeline twice: PASS 0 packs (F m) + i G
// and writes the slab partials (sum of (F m)^2 / w, the maxima of F and m),
// PASS 1 packs m + i G^2 (masked_nsqecc_ops.jl:19-42).
template <int L, int ALG, typename T, int LPL, int PASS = 0>
__global__ __launch_bounds__((Cfg3<L, T, LPL>::TPL)) void slab_fwd_kernel(Job J, Scratch3 sc, int64_t ia_base,
                                                                          int64_t n_ia, const double* sat1_img1) {
  constexpr bool kMask = ALG == ALG_MASK_NSQECC;
  static_assert(PASS == 0 || kMask, "a second pass is the masked measure's");
  using C = Cfg3<L, T, LPL>;
  constexpr int M = C::M, LD = C::LD;
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* Z = reinterpret_cast<cpx<T>*>(smem);
  double* red = reinterpret_cast<double*>(smem + (size_t)L * LD * sizeof(cpx<T>));
  const int t = threadIdx.x;
  const int line = t / LPL, q = t % LPL;
  const int64_t il = blockIdx.x / L;  // IA within the chunk
  const int i2 = blockIdx.x % L;
  const int64_t ia = ia_base + il;
  if (il >= n_ia) return;
  const Geo3 G = geo3_of(J, ia);
  const Geo &g0 = G.g[0], &g1 = G.g[1], &g2 = G.g[2];
  const int dt = J.in_dtype;
  const int64_t d0 = J.dims[0], d01 = J.dims[0] * J.dims[1];

  Means3 mu{0.f, 0.f, 0.0, 0.0};
  if constexpr (ALG == ALG_ZNCC) mu = means3(J, G, sat1_img1, ia);

  const int k0 = line;
  const bool f_pl = i2 < J.is[2];
  const bool g_pl = i2 >= g2.goff && i2 < g2.goff + g2.gsz;
  const bool f_row = f_pl && k0 < J.is[0];
  const bool g_row = g_pl && k0 >= g0.goff && k0 < g0.goff + g0.gsz;
  const int64_t f_base = (g0.tlf + k0) + d0 * g1.tlf + d01 * (g2.tlf + i2);
  const int64_t g_base = (g0.sr + (k0 - g0.goff)) + d0 * (g1.sr - g1.goff) + d01 * (g2.sr + (i2 - g2.goff));
  const bool want_max = J.has_threshold || J.has_mask_threshold;
  double sumF2 = 0.0, mx1 = -INFINITY, mx2 = -INFINITY;
  cpx<T> x[M];
  constexpr bool kBox = ALG == ALG_ZNCC || ALG == ALG_NSQECC;
  double gq[kBox ? M : 1];  // G^2 (ZNCC: G'^2) of the lane's samples, 0 outside the data
  // Every lane issues all its loads before using any: positions outside the
  // IA / clipped search region read an in-volume voxel instead (the IA origin,
  // or the search region's first column in this lane's row and plane) and
  // their values are discarded, so no load is predicated.
  auto f_at = [&](int k1) { return f_row && k1 < J.is[1]; };
  auto g_at = [&](int k1) { return g_row && k1 >= g1.goff && k1 < g1.goff + g1.gsz; };
  const int64_t f_lane = f_row ? f_base : (g0.tlf + d0 * g1.tlf + d01 * g2.tlf);
  const int64_t g_lane = g_row ? g_base : (g0.sr + d0 * (g1.sr - g1.goff) + d01 * g2.sr);
  auto load = [&](auto tin) {
    using TIN = decltype(tin);
    const gptr<TIN> pf = as_global<TIN>(J.img1) + f_lane;
    const gptr<TIN> pg = as_global<TIN>(J.img2) + g_lane;
    const gptr<TIN> pm = as_global<TIN>(kMask ? J.mask : J.img1) + f_lane;
    TIN fv[M], gv[M], mv[kMask ? M : 1];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int k1 = LPL * k + q;
      if (PASS == 0) fv[k] = pf[f_at(k1) ? d0 * k1 : 0];
      if constexpr (kMask) mv[k] = pm[f_at(k1) ? d0 * k1 : 0];
      gv[k] = pg[d0 * (g_at(k1) ? k1 : g1.goff)];
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int k1 = LPL * k + q;
      double a = 0.0, g = 0.0;
      if (f_at(k1)) {
        if constexpr (kMask) {
          const double m = (double)mv[k];
          if constexpr (PASS == 0) {
            const TIN f = fv[k];
            if (want_max) {
              mx1 = fmax(mx1, (double)f);
              mx2 = fmax(mx2, m);
            }
            a = (double)f * m;  // masked_nsqecc_ops.jl:19-25
            const double wi = m > 0 ? m : 1.0;
            sumF2 += a * a / wi;  // mask_nsqecc.jl:144-149
          } else {
            a = m;
          }
        } else {
          const TIN f = fv[k];
          if (want_max) mx1 = fmax(mx1, (double)f);
          if constexpr (ALG == ALG_ZNCC)
            a = std::is_same<TIN, float>::value ? (double)((float)f - mu.f32) : (double)f - mu.f64;
          else
            a = (double)f;
          if constexpr (ALG != ALG_FFTCC) sumF2 += a * a;
        }
      }
      if (g_at(k1)) {
        if constexpr (ALG == ALG_ZNCC)
          g = std::is_same<TIN, float>::value ? (double)((float)gv[k] - mu.g32) : (double)gv[k] - mu.g64;
        else
          g = (double)gv[k];
        if constexpr (PASS == 1) g = g * g;  // masked_nsqecc_ops.jl:36-42
      }
      if constexpr (kBox) gq[k] = g * g;
      x[k] = cpx<T>{(T)a, (T)g};
    }
  };
  if constexpr (kBox) {
#pragma unroll
    for (int k = 0; k < M; ++k) gq[k] = 0.0;
  }
  if (line < L) {
    if (dt == DT_F32) load(float{});
    else load(double{});
  }
  if constexpr (PASS == 0) {
    sumF2 = Group<C::TPL>::reduce(sumF2, OpAdd{}, red, t);
    mx1 = Group<C::TPL>::reduce(mx1, OpMax{}, red, t);
    if constexpr (kMask) mx2 = Group<C::TPL>::reduce(mx2, OpMax{}, red, t);
    if (t == 0) {
      double* part = sc.part + il * (3 * L);
      part[i2] = sumF2;
      part[L + i2] = mx1;
      part[2 * L + i2] = mx2;
    }
  }

  // forward along dim 2 (registers), then dim 1 through LDS
  const int cM = (LPL == 1 ? 0 : chunk_of<LPL>(q)) * M;
  if (line < L) {
    line_fwd<L, LPL, T>(x, q);
#pragma unroll
    for (int k = 0; k < M; ++k) Z[(cM + k) * LD + line] = x[k];
  }
  __syncthreads();
  if (line < L) {
#pragma unroll
    for (int k = 0; k < M; ++k) x[k] = Z[line * LD + LPL * k + q];
    line_fwd<L, LPL, T>(x, q);
  }
  // plane layout in scratch: [i2][k1][k0].  A lane holds M consecutive k0 of
  // its row k1, so storing them directly would scatter every store over 64
  // rows; the plane goes through LDS densely instead (k0 XOR (k1 & 3), which
  // spreads the lanes of a write over the banks) and leaves in whole 16-byte
  // vectors, consecutive lanes on consecutive addresses.
  static_assert(L % 4 == 0 && (L * L * sizeof(cpx<T>)) % 16 == 0, "16-byte vectors over whole groups of 4");
  __syncthreads();  // every lane has read Z
  if (line < L) {
#pragma unroll
    for (int k = 0; k < M; ++k) Z[line * L + ((cM + k) ^ (line & 3))] = x[k];
  }
  __syncthreads();
  {
    cpx<T>* Wp = (cpx<T>*)sc.w + il * sc.w_stride + (int64_t)i2 * L * L;
    constexpr int EPV = 16 / (int)sizeof(cpx<T>);  // complex values per vector: 2 (FP32) or 1 (FP64)
    constexpr int NV = L * L / EPV;
    for (int v = t; v < NV; v += C::TPL) {
      const int e = v * EPV;  // first element: row k1 = e / L, column k0 = e % L
      const int r = e / L, c0 = e - r * L;
      const int sw = r & 3;
      if constexpr (EPV == 2) {
        // elements c0 and c0 + 1 (c0 even) sit in one aligned pair at
        // c0 ^ (sw & 2), swapped when sw is odd
        const uint4 raw = *reinterpret_cast<const uint4*>(&Z[r * L + (c0 ^ (sw & 2))]);
        *reinterpret_cast<uint4*>(&Wp[e]) = (sw & 1) ? uint4{raw.z, raw.w, raw.x, raw.y} : raw;
      } else {
        Wp[e] = Z[r * L + (c0 ^ sw)];
      }
    }
  }
  if constexpr (kBox) {
    // The slab's window sums of G^2 (ZNCC: of G'^2, the rounded differences, as
    // zncc_ops.jl:22-24 and integralArraySQ!, common_ops.jl:142-180): an
    // inclusive summed-area table of the slab, P(a, b) = sum over rows <= a and
    // columns <= b, built in the transform buffer once the transform has left
    // it (L x L doubles fit in L x (L + 1) complex values of either type), then
    // for every window position (s0, s1) the box [s0, s0 + is0) x [s1, s1 + is1)
    // in the order of the 2-D window sums (common_ops.jl:203-245).  Positions
    // outside the clipped search region hold 0, so these are the sums of the
    // clipped windows.  Each prefix pass splits a line over PS adjacent lanes:
    // every lane scans its segment in registers and adds the totals of the
    // line's earlier segments.
    // (columns of the table at stride L + 1: the prefix passes' lanes then
    // fall on distinct LDS banks)
    constexpr int PL = L + 1;
    static_assert((size_t)L * PL * sizeof(double) <= (size_t)L * LD * sizeof(cpx<T>), "table fits the buffer");
    static_assert(C::TPL % L == 0, "whole lanes per line");
    constexpr int PS = C::TPL / L, RS = L / PS;
    static_assert(RS * PS == L && (PS & (PS - 1)) == 0, "equal segments, a power of two of lanes");
    double* P = reinterpret_cast<double*>(smem);
    __syncthreads();  // every lane has read the buffer
    if (line < L) {
#pragma unroll
      for (int k = 0; k < M; ++k) P[line + PL * (LPL * k + q)] = gq[k];
    }
    __syncthreads();
    const int seg = t % PS, ln = t / PS;
    auto pass = [&](int line_stride, int elem_stride) {
      double v[RS];
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < RS; ++k) {
        acc += P[ln * line_stride + (seg * RS + k) * elem_stride];
        v[k] = acc;
      }
      double inc = acc;
#pragma unroll
      for (int d = 1; d < PS; d <<= 1) {
        const double o = __shfl_up(inc, d, PS);
        if (seg >= d) inc += o;
      }
      const double off = inc - acc;
#pragma unroll
      for (int k = 0; k < RS; ++k) P[ln * line_stride + (seg * RS + k) * elem_stride] = v[k] + off;
      __syncthreads();
    };
    pass(PL, 1);  // down each column (dim 1)
    pass(1, PL);  // along each row (dim 2)
    const int ns0 = 2 * J.sm[0] + 1, ns1 = 2 * J.sm[1] + 1;
    const int is0 = J.is[0], is1 = J.is[1];
    auto at = [&](int a, int bb) { return (a < 0 || bb < 0) ? 0.0 : P[a + PL * bb]; };
    double* B = sc.box + il * sc.box_stride + (int64_t)i2 * ns0 * ns1;
    for (int e = t; e < ns0 * ns1; e += C::TPL) {
      const int s0 = e % ns0, s1 = e / ns0;
      B[e] = ((at(s0 + is0 - 1, s1 + is1 - 1) - at(s0 - 1, s1 + is1 - 1)) - at(s0 + is0 - 1, s1 - 1)) +
             at(s0 - 1, s1 - 1);
    }
  }
}

// ---------------------------------------------------------------------------
// K2: along dim 3 for the planes k1 and -k1, conjugate product, inverse.
template <int L, typename T, int LPL>
__global__ __launch_bounds__((Cfg3<L, T, LPL>::NT2)) void zline_kernel(Scratch3 sc, int64_t n_ia, int ns2) {
  using C = Cfg3<L, T, LPL>;
  constexpr int M = C::M, LD = C::LD;
  constexpr int NP = L / 2 + 1;  // plane pairs k1 in [0, L/2]
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* Pl = reinterpret_cast<cpx<T>*>(smem);  // [plane 0/1][k2][k0], LD stride
  const int t = threadIdx.x;
  const int64_t il = blockIdx.x / NP;
  const int ka = blockIdx.x % NP;
  if (il >= n_ia) return;
  const int kb = (L - ka) % L;
  const int pl = t / (L * LPL);  // which plane this lane works on
  const int tt = t % (L * LPL);
  const int line = tt / LPL, q = tt % LPL;  // line = k0
  const int k1 = pl == 0 ? ka : kb;
  const bool lane = pl < 2 && !(pl == 1 && ka == kb);
  cpx<T>* W = (cpx<T>*)sc.w + il * sc.w_stride;
  const int cM = (LPL == 1 ? 0 : chunk_of<LPL>(q)) * M;
  cpx<T> x[M];
  if (lane) {
#pragma unroll
    for (int k = 0; k < M; ++k) x[k] = W[(int64_t)(LPL * k + q) * L * L + (int64_t)k1 * L + line];
    line_fwd<L, LPL, T>(x, q);
#pragma unroll
    for (int k = 0; k < M; ++k) Pl[(pl * L + cM + k) * LD + line] = x[k];
  }
  __syncthreads();
  // The correlation is real: plane L - ka of the product is the reflected
  // conjugate of plane ka, so only plane ka (<= L/2) is multiplied and
  // inverted; plane kb was transformed above as its partner.
  const bool keep = lane && pl == 0;
  if (keep) {
    // partner of (k0, k1, k2) is (-k0, -k1, -k2): plane kb
    const int pp = (ka == kb) ? 0 : 1;
    const int k0n = (L - line) % L;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int k2 = cM + k;
      const int k2n = (L - k2) % L;
      cpx<T> z = x[k];
      cpx<T> n = Pl[(pp * L + k2n) * LD + k0n];
      T ar = T(0.5) * (z.x + n.x), ai = T(0.5) * (z.y - n.y);
      T br = T(0.5) * (z.y + n.y), bi = T(0.5) * (n.x - z.x);
      x[k] = cpx<T>{ar * br + ai * bi, ar * bi - ai * br};
    }
    line_inv<L, LPL, T>(x, q);
  }
  __syncthreads();  // every forward value of these planes is read before any write
  if (keep) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int s2 = LPL * m + q;
      if (s2 < ns2) W[(int64_t)s2 * L * L + (int64_t)k1 * L + line] = x[m];
    }
  }
}

// ---------------------------------------------------------------------------
// K3: inverse 2-D transform of plane s2 down to the (ns0 x ns1) window.
template <int L, typename T, int LPL>
__global__ __launch_bounds__((Cfg3<L, T, LPL>::TPL)) void slab_inv_kernel(Scratch3 sc, int64_t n_ia, int ns0,
                                                                          int ns1, int ns2) {
  using C = Cfg3<L, T, LPL>;
  constexpr int M = C::M, LD = C::LD;
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* Z = reinterpret_cast<cpx<T>*>(smem);
  const int t = threadIdx.x;
  const int line = t / LPL, q = t % LPL;
  const int64_t il = blockIdx.x / ns2;
  const int s2 = blockIdx.x % ns2;
  if (il >= n_ia) return;
  const cpx<T>* W = (const cpx<T>*)sc.w + il * sc.w_stride + (int64_t)s2 * L * L;
  const int cM = (LPL == 1 ? 0 : chunk_of<LPL>(q)) * M;
  cpx<T> x[M];
  // Each plane's spectrum is Hermitian in (k0, k1) (the correlation is real)
  // and only columns k1 <= L/2 were written: inverse along dim 1 for those,
  // then along dim 2 with two real rows packed per complex transform and the
  // columns above L/2 rebuilt as conjugates.
  constexpr int KH = L / 2 + 1;
  if (line < KH) {
#pragma unroll
    for (int k = 0; k < M; ++k) x[k] = W[(int64_t)line * L + cM + k];
    line_inv<L, LPL, T>(x, q);
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int s0 = LPL * m + q;
      if (s0 < ns0) Z[line * LD + s0] = x[m];
    }
  }
  __syncthreads();
  const int npair = (ns0 + 1) / 2;
  if (line < npair) {
    const int sa = 2 * line, sb = 2 * line + 1;
    const bool has_b = sb < ns0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int k1 = cM + k;
      const bool mirror = k1 >= KH;
      const int kc = mirror ? L - k1 : k1;
      const cpx<T> ya = Z[kc * LD + sa];
      const cpx<T> yb = has_b ? Z[kc * LD + sb] : cpx<T>{T(0), T(0)};
      const T ayi = mirror ? -ya.y : ya.y, byi = mirror ? -yb.y : yb.y;
      x[k] = cpx<T>{ya.x - byi, ayi + yb.x};
    }
    line_inv<L, LPL, T>(x, q);
    T* S = (T*)sc.s + il * sc.s_stride + (int64_t)s2 * ns0 * ns1;
    const T kScale = T(1) / T(L * L * L);
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int s1 = LPL * m + q;
      if (s1 < ns1) {
        S[s1 * ns0 + sa] = x[m].x * kScale;
        if (has_b) S[s1 * ns0 + sb] = x[m].y * kScale;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// K4: normalisation, peak, subpixel, SN and output for one IA.
template <int ALG, typename T>
__global__ __launch_bounds__(256) void finalize3d_kernel(Job J, Scratch3 sc, int64_t ia_base, int64_t n_ia, int L,
                                                         const double* sat1_img1) {
  constexpr int NT = 256;
  __shared__ double red[32];
  const int t = threadIdx.x;
  const int64_t il = blockIdx.x;
  if (il >= n_ia) return;
  const int64_t ia = ia_base + il;
  const Geo3 G = geo3_of(J, ia);
  const Geo &g0 = G.g[0], &g1 = G.g[1], &g2 = G.g[2];
  const int sm0 = J.sm[0], sm1 = J.sm[1], sm2 = J.sm[2];
  const int is0 = J.is[0], is1 = J.is[1], is2 = J.is[2];
  const int ns0 = 2 * sm0 + 1, ns1 = 2 * sm1 + 1, ns2 = 2 * sm2 + 1;
  const int nS = ns0 * ns1 * ns2;
  T* S = (T*)sc.s + il * sc.s_stride;

  // IA constants from the per-slab partials
  double sumF2 = 0.0, mx1 = -INFINITY, mx2 = -INFINITY;
  const double* part = sc.part + il * (3 * L);
#pragma unroll 8
  for (int i = 0; i < L; ++i) {
    sumF2 += part[i];
    mx1 = fmax(mx1, part[L + i]);
    mx2 = fmax(mx2, part[2 * L + i]);
  }
  // (mx2, the mask maximum, is only gathered by the masked measure, which is
  // the only one the host gives a mask threshold)
  const bool skip = J.skip ? (J.skip[ia] != 0)
                           : ((J.has_threshold && mx1 < J.threshold) ||
                              (J.has_mask_threshold && mx2 < J.mask_threshold));

  if constexpr (ALG == ALG_MASK_NSQECC) {
    // mask_nsqecc.jl:155 and masked_nsqecc_ops.jl:123-143: C1 = (F m) x G,
    // C2 = m x G^2, num = sum(F^2) + C2 - 2 C1, den = sqrt|C2| sqrt(sum(F^2))
    const double sqF = sqrt(sumF2);
    const T* S2 = (const T*)sc.s2 + il * sc.s_stride;
    for (int e = t; e < nS; e += NT) {
      const double c1 = (double)S[e], c2 = (double)S2[e];
      const double num = (-2.0 * c1 + sumF2) + c2;
      const double den = sqrt(fabs(c2)) * sqF;
      S[e] = (T)(den / (num + den));
    }
    __syncthreads();
  }

  if constexpr (ALG == ALG_ZNCC || ALG == ALG_NSQECC) {
    // W(s) = sum over the window's planes s2 .. s2 + is2 - 1 of the slabs' window
    // sums at (s0, s1), as the difference of two running sums along dim 3 (the
    // prefix-sum form of the integral array's dim 3); one thread per (s0, s1)
    const double sqF = sqrt(sumF2);
    const int nP = ns0 * ns1;
    const double* B = sc.box + il * sc.box_stride;
    for (int e = t; e < nP; e += NT) {
      double hi = 0.0, lo = 0.0;  // sums over planes < s2 + is2 and < s2
      // (unrolled: the loads of a group are in flight together, in the order
      // of the sum)
#pragma unroll 8
      for (int k = 0; k < is2; ++k) hi += B[(int64_t)k * nP + e];
#pragma unroll 4
      for (int s2 = 0; s2 < ns2; ++s2) {
        if (s2 > 0) {
          hi += B[(int64_t)(s2 + is2 - 1) * nP + e];
          lo += B[(int64_t)(s2 - 1) * nP + e];
        }
        double W = hi - lo;
        W = W > 0.0 ? W : 0.0;
        const int64_t li = e + (int64_t)nP * s2;
        const double c = (double)S[li];
        const double den = sqrt(W) * sqF;
        double r;
        if constexpr (ALG == ALG_ZNCC) {
          r = c / den;  // zncc_ops.jl:39-48
        } else {
          const double num = (sumF2 - 2.0 * c) + W;  // nsqecc_ops.jl:58-60
          r = den / (num + den);                     // nsqecc_ops.jl:72-73
        }
        S[li] = (T)r;
      }
    }
    __syncthreads();
  }

  // peak over [off_lo, 2SM - off_hi] per dim, column-major first strict max
  const int lo0 = g0.goff, hi0 = 2 * sm0 - g0.off_hi;
  const int lo1 = g1.goff, hi1 = 2 * sm1 - g1.off_hi;
  const int lo2 = g2.goff, hi2 = 2 * sm2 - g2.off_hi;
  const int w0 = hi0 - lo0 + 1, w1 = hi1 - lo1 + 1, w2 = hi2 - lo2 + 1;
  const int nwin = (w0 > 0 && w1 > 0 && w2 > 0) ? w0 * w1 * w2 : 0;
  Best<T> bst{T(0), INT32_MAX};
  BoxWalk pw(t, NT, w0, w1);
#pragma unroll 4
  for (int e = t; e < nwin; e += NT, pw.step()) {
    const int p0 = lo0 + pw.a, p1 = lo1 + pw.b, p2 = lo2 + pw.c;
    const int li = p0 + ns0 * (p1 + ns1 * p2);
    const T v = S[li];
    if (v > bst.v) bst = Best<T>{v, li};
  }
  Group<NT>::steps([&](auto mv) { bst = best_of(bst, Best<T>{mv(T(0), bst.v), mv((int32_t)INT32_MAX, bst.i)}); });
  {
    T* rv = reinterpret_cast<T*>(red);
    int32_t* ri = reinterpret_cast<int32_t*>(red + 16);
    __syncthreads();
    if ((t & 63) == 63) {
      rv[t >> 6] = bst.v;
      ri[t >> 6] = bst.i;
    }
    __syncthreads();
    Best<T> r{rv[0], ri[0]};
    for (int w = 1; w < NT / 64; ++w) r = best_of(r, Best<T>{rv[w], ri[w]});
    bst = r;
  }

  // SN over [off_lo, 2SM - 1 - off_hi] per dim (natural 3-D extension)
  double sn = 0.0;
  if (J.compute_sn) {
    const int u0 = (2 * sm0 - 1 - g0.off_hi) - lo0 + 1;
    const int u1 = (2 * sm1 - 1 - g1.off_hi) - lo1 + 1;
    const int u2 = (2 * sm2 - 1 - g2.off_hi) - lo2 + 1;
    const int nsn = (u0 > 0 && u1 > 0 && u2 > 0) ? u0 * u1 * u2 : 0;
    double mx = -INFINITY, sm = 0.0;
    BoxWalk sw(t, NT, u0, u1);
#pragma unroll 4
    for (int e = t; e < nsn; e += NT, sw.step()) {
      const int p0 = lo0 + sw.a, p1 = lo1 + sw.b, p2 = lo2 + sw.c;
      const double v = (double)S[p0 + ns0 * (p1 + ns1 * p2)];
      mx = (v > mx || v != v) ? v : mx;
      sm += v;
    }
    mx = Group<NT>::reduce(mx, OpMaxNaN{}, red, t);
    sm = Group<NT>::reduce(sm, OpAdd{}, red, t);
    sn = nsn > 0 ? mx / ((sm - mx) / (double)(nsn - 1)) : (double)NAN;
  }

  if (t == 0) {
    double r[3];
    const int smv[3] = {sm0, sm1, sm2};
    if (skip) {
      r[0] = r[1] = r[2] = 0.0;
      sn = 0.0;
    } else if (bst.i == INT32_MAX) {
      for (int d = 0; d < 3; ++d) r[d] = -(double)(smv[d] + 1);
    } else {
      const int p[3] = {bst.i % ns0, (bst.i / ns0) % ns1, bst.i / (ns0 * ns1)};
      const bool interior = p[0] > lo0 && p[0] < hi0 && p[1] > lo1 && p[1] < hi1 && p[2] > lo2 && p[2] < hi2;
      for (int d = 0; d < 3; ++d) r[d] = (double)(p[d] + 1) - (double)(smv[d] + 1);
      if (interior) {
        auto at = [&](int a, int b, int c) { return (double)S[(p[0] + a) + ns0 * ((p[1] + b) + ns1 * (p[2] + c))]; };
        double mn = INFINITY;
        bool nan = false;
        for (int c = -1; c <= 1; ++c)
          for (int b = -1; b <= 1; ++b)
            for (int a = -1; a <= 1; ++a) {
              const double v = at(a, b, c);
              if (v != v) nan = true;
              mn = v < mn ? v : mn;
            }
        if (nan) mn = NAN;
        const double lmx = log(1.0 + at(0, 0, 0) - mn);
        const double lm[3] = {log(1.0 + at(-1, 0, 0) - mn), log(1.0 + at(0, -1, 0) - mn), log(1.0 + at(0, 0, -1) - mn)};
        const double lp[3] = {log(1.0 + at(1, 0, 0) - mn), log(1.0 + at(0, 1, 0) - mn), log(1.0 + at(0, 0, 1) - mn)};
        for (int d = 0; d < 3; ++d) {
          const double qd = (lm[d] - lp[d]) / (2 * lm[d] - 4 * lmx + 2 * lp[d]);
          r[d] = ((double)(p[d] + 1) + qd) - (double)(smv[d] + 1);
        }
      }
    }
    if (!skip)  // the displacement is measured from the shifted search region
      for (int d = 0; d < 3; ++d) r[d] += (double)G.d[d];
    if (J.peaks) {
      const bool has = !skip && bst.i != INT32_MAX;
      J.peaks[3 * ia + 0] = has ? bst.i % ns0 + 1 : 0;
      J.peaks[3 * ia + 1] = has ? (bst.i / ns0) % ns1 + 1 : 0;
      J.peaks[3 * ia + 2] = has ? bst.i / (ns0 * ns1) + 1 : 0;
    }
    if (J.out_dtype == DT_F32) {
      float* vf = (float*)J.vf;
      for (int d = 0; d < 3; ++d) vf[3 * ia + d] = (float)r[d];
      if (J.sn) ((float*)J.sn)[ia] = (float)sn;
    } else {
      double* vf = (double*)J.vf;
      for (int d = 0; d < 3; ++d) vf[3 * ia + d] = r[d];
      if (J.sn) ((double*)J.sn)[ia] = sn;
    }
  }
}

// ---------------------------------------------------------------------------
// Host-side launch of one chunk of IAs [ia_base, ia_base + n).
template <int L, typename T, int LPL>
size_t scratch3_per_ia(int alg, int ns0, int ns1, int ns2) {
  size_t w = (size_t)L * L * L * sizeof(cpx<T>);
  size_t s = ((size_t)ns0 * ns1 * ns2 * sizeof(T) + 255) & ~(size_t)255;
  if (alg == ALG_MASK_NSQECC) s *= 2;  // the second correlation's window
  size_t p = (size_t)3 * L * sizeof(double);
  size_t bx = (alg == ALG_ZNCC || alg == ALG_NSQECC) ? (size_t)L * ns0 * ns1 * sizeof(double) : 0;
  return w + s + p + bx;
}

template <int L, typename T, int LPL>
hipError_t launch3d_chunk(const Job& J, void* scratch, int64_t chunk_cap, int64_t ia_base, int64_t n,
                          const double* sat1_img1, hipStream_t s) {
  using C = Cfg3<L, T, LPL>;
  const int ns0 = 2 * J.sm[0] + 1, ns1 = 2 * J.sm[1] + 1, ns2 = 2 * J.sm[2] + 1;
  Scratch3 sc;
  sc.w_stride = (int64_t)L * L * L;
  sc.s_stride = (int64_t)ns0 * ns1 * ns2;
  char* b = (char*)scratch;
  sc.w = b;
  b += (size_t)chunk_cap * L * L * L * sizeof(cpx<T>);
  sc.s = b;
  b += (((size_t)chunk_cap * ns0 * ns1 * ns2 * sizeof(T)) + 255) & ~(size_t)255;
  sc.s2 = nullptr;
  if (J.alg == ALG_MASK_NSQECC) {
    sc.s2 = b;
    b += (((size_t)chunk_cap * ns0 * ns1 * ns2 * sizeof(T)) + 255) & ~(size_t)255;
  }
  sc.part = (double*)b;
  b += (size_t)chunk_cap * 3 * L * sizeof(double);
  const bool box = J.alg == ALG_ZNCC || J.alg == ALG_NSQECC;
  sc.box = box ? (double*)b : nullptr;
  sc.box_stride = (int64_t)L * ns0 * ns1;
  const size_t lds1 = (size_t)L * C::LD * sizeof(cpx<T>) + 16 * sizeof(double);
  const size_t lds2 = (size_t)2 * L * C::LD * sizeof(cpx<T>);
  const size_t lds3 = (size_t)L * C::LD * sizeof(cpx<T>);
  allow_full_lds((const void*)slab_fwd_kernel<L, ALG_FFTCC, T, LPL>);
  allow_full_lds((const void*)slab_fwd_kernel<L, ALG_ZNCC, T, LPL>);
  allow_full_lds((const void*)slab_fwd_kernel<L, ALG_NSQECC, T, LPL>);
  allow_full_lds((const void*)slab_fwd_kernel<L, ALG_MASK_NSQECC, T, LPL, 0>);
  allow_full_lds((const void*)slab_fwd_kernel<L, ALG_MASK_NSQECC, T, LPL, 1>);
  allow_full_lds((const void*)zline_kernel<L, T, LPL>);
  allow_full_lds((const void*)slab_inv_kernel<L, T, LPL>);
  switch (J.alg) {
    case ALG_FFTCC:
      hipLaunchKernelGGL((slab_fwd_kernel<L, ALG_FFTCC, T, LPL>), dim3((unsigned)(n * L)), dim3(C::TPL), lds1, s, J,
                         sc, ia_base, n, sat1_img1);
      break;
    case ALG_ZNCC:
      hipLaunchKernelGGL((slab_fwd_kernel<L, ALG_ZNCC, T, LPL>), dim3((unsigned)(n * L)), dim3(C::TPL), lds1, s, J,
                         sc, ia_base, n, sat1_img1);
      break;
    case ALG_NSQECC:
      hipLaunchKernelGGL((slab_fwd_kernel<L, ALG_NSQECC, T, LPL>), dim3((unsigned)(n * L)), dim3(C::TPL), lds1, s,
                         J, sc, ia_base, n, sat1_img1);
      break;
    case ALG_MASK_NSQECC: {
      // the second correlation first, into S2 (K3 writes the window its
      // Scratch3 names), then the first, into S
      Scratch3 sc2 = sc;
      sc2.s = sc.s2;
      hipLaunchKernelGGL((slab_fwd_kernel<L, ALG_MASK_NSQECC, T, LPL, 1>), dim3((unsigned)(n * L)), dim3(C::TPL), lds1,
                         s, J, sc2, ia_base, n, sat1_img1);
      hipLaunchKernelGGL((zline_kernel<L, T, LPL>), dim3((unsigned)(n * (L / 2 + 1))), dim3(C::NT2), lds2, s, sc2, n,
                         ns2);
      hipLaunchKernelGGL((slab_inv_kernel<L, T, LPL>), dim3((unsigned)(n * ns2)), dim3(C::TPL), lds3, s, sc2, n, ns0,
                         ns1, ns2);
      hipLaunchKernelGGL((slab_fwd_kernel<L, ALG_MASK_NSQECC, T, LPL, 0>), dim3((unsigned)(n * L)), dim3(C::TPL), lds1,
                         s, J, sc, ia_base, n, sat1_img1);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL((zline_kernel<L, T, LPL>), dim3((unsigned)(n * (L / 2 + 1))), dim3(C::NT2), lds2, s, sc, n,
                     ns2);
  hipLaunchKernelGGL((slab_inv_kernel<L, T, LPL>), dim3((unsigned)(n * ns2)), dim3(C::TPL), lds3, s, sc, n, ns0, ns1,
                     ns2);
  switch (J.alg) {
    case ALG_FFTCC:
      hipLaunchKernelGGL((finalize3d_kernel<ALG_FFTCC, T>), dim3((unsigned)n), dim3(256), 0, s, J, sc, ia_base, n, L,
                         sat1_img1);
      break;
    case ALG_ZNCC:
      hipLaunchKernelGGL((finalize3d_kernel<ALG_ZNCC, T>), dim3((unsigned)n), dim3(256), 0, s, J, sc, ia_base, n, L,
                         sat1_img1);
      break;
    case ALG_NSQECC:
      hipLaunchKernelGGL((finalize3d_kernel<ALG_NSQECC, T>), dim3((unsigned)n), dim3(256), 0, s, J, sc, ia_base, n,
                         L, sat1_img1);
      break;
    case ALG_MASK_NSQECC:
      hipLaunchKernelGGL((finalize3d_kernel<ALG_MASK_NSQECC, T>), dim3((unsigned)n), dim3(256), 0, s, J, sc, ia_base,
                         n, L, sat1_img1);
      break;
  }
  return hipGetLastError();
}

}  // namespace mqpiv
