// mqpiv_kernels2d.hpp -- fused 2-D cross-correlation PIV kernel for gfx950.
//
// One thread group processes one interrogation area (IA) end to end:
//   gather IA and search window from HBM (coalesced along dim 1)
//   -> pack Z = F + iG in LDS -> line FFTs along dim 1 then dim 2 in VGPRs
//   -> P = conj(F^) G^ (partner bins read back through LDS)
//   -> inverse line FFTs (dim 2, then dim 1 only for the 2SM+1 needed columns)
//   -> FP64 box sums of G^2 for the ZNCC/NSQECC denominators
//   -> similarity on the fully-overlapping window -> argmax -> Gaussian
//   subpixel -> SN -> VF/SN in the reference's layout.
// T is the transform's scalar type (float: FP32 transforms with FP64 sums and
// denominators; double: every step in FP64).  LPL lanes share one line.
//
// Reference behaviour restated (files relative to the reference repo root):
//   coordinates/offsets     src/utils/piv_ops.jl:38-85 (incl. SR_TLF_off, :72)
//   padded copies           src/utils/piv_ops.jl:99-127, zncc_ops.jl:3-25,
//                           masked_nsqecc_ops.jl:19-42
//   correlation             src/crosscorrelation_algorithms/fftcc.jl:101-126,
//                           src/utils/fft_ops.jl:95-105
//   ZNCC / NSQECC / masked  zncc_ops.jl:29-51, nsqecc_ops.jl:1-26,
//                           masked_nsqecc_ops.jl:123-143, mask_nsqecc.jl:141-160
//   window sums             src/utils/common_ops.jl:142-245
//   peak                    src/utils/common_ops.jl:98-124
//   subpixel                src/utils/piv_ops.jl:249-326
//   SN                      src/utils/piv_ops.jl:333-367
//   skip                    src/utils/piv_ops.jl:134-157 (filtFun = maxval)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "mqpiv_fft.hpp"
#include "mqpiv_internal.h"
#include "mqpiv_lpl2d.h"

namespace mqpiv {

template <int L, typename T, int LPL>
struct Cfg2 {
  static constexpr int H = L / LPL;                       // samples per lane
  static constexpr int LANES = L * LPL;                   // lanes carrying lines
  static constexpr int TPI = LANES <= 32 ? LANES : ((LANES + 63) / 64) * 64;  // lanes per IA
  static constexpr int IPB = LANES <= 32 ? 64 / LANES : 1;                   // IAs per block
  static constexpr int NT = TPI * IPB;                    // threads per block
  static constexpr int LD = L + 1;                        // LDS row stride (complex)
  static constexpr int NW = (TPI + 63) / 64;              // waves per IA
  static_assert(NW <= 16, "reduction slots hold 16 waves");
  static_assert(LPL == 1 || LPL == 2 || LPL == 4 || LPL == 8, "1, 2, 4 or 8 lanes per line");
};

// Global-address-space view of a device pointer: loads through it are
// global_load (vector memory only) rather than flat loads, which also count
// against the LDS counter and force waits on both.
template <typename T>
using gptr = const T __attribute__((address_space(1)))*;
template <typename T>
__device__ __forceinline__ gptr<T> as_global(const void* p) {
  return (gptr<T>)p;
}
// Writable global-address-space pointer (the global part of a transform buffer
// that does not fit in LDS).
template <typename T>
using gmut = T __attribute__((address_space(1)))*;
// The transform buffer of one IA, addressed by (column, row) with row stride
// L + 1: in LDS (ZLds) or, for the lengths whose buffer does not fit (gz_len),
// in global memory (ZGlob).
template <typename T>
using vec2_t = T __attribute__((ext_vector_type(2)));
template <typename T, int LD>
struct ZLds {
  cpx<T>* z;
  __device__ __forceinline__ cpx<T> ld(int col, int row) const { return z[col * LD + row]; }
  __device__ __forceinline__ void st(int col, int row, cpx<T> v) const { z[col * LD + row] = v; }
  __device__ __forceinline__ T* real() const { return reinterpret_cast<T*>(z); }
};
template <typename T, int LD>
struct ZGlob {
  gmut<cpx<T>> g;  // one vector access per value
  __device__ __forceinline__ cpx<T> ld(int col, int row) const {
    const vec2_t<T> v = ((gmut<vec2_t<T>>)g)[col * LD + row];
    return cpx<T>{v.x, v.y};
  }
  __device__ __forceinline__ void st(int col, int row, cpx<T> v) const {
    vec2_t<T> w;
    w.x = v.x;
    w.y = v.y;
    ((gmut<vec2_t<T>>)g)[col * LD + row] = w;
  }
  __device__ __forceinline__ gmut<T> real() const { return (gmut<T>)g; }
};

__device__ __forceinline__ double load_px(const void* img, int dt, int64_t i) {
  return dt == DT_F32 ? (double)((const float*)img)[i] : ((const double*)img)[i];
}

// Cross-lane moves for reductions.  A mover returns, for every lane, the value
// of the lane it is paired with in one reduction step; where a step has no
// source lane the lane receives `id` (the reduction's identity).  32-bit DPP
// moves (VALU) do the work; doubles move as two halves.
template <int CTRL, int ROWM, int BANKM>
__device__ __forceinline__ int dpp_i32(int id, int v) {
  return __builtin_amdgcn_update_dpp(id, v, CTRL, ROWM, BANKM, false);
}
template <int CTRL, int ROWM, int BANKM>
struct DppMove {
  __device__ __forceinline__ int operator()(int id, int v) const { return dpp_i32<CTRL, ROWM, BANKM>(id, v); }
  __device__ __forceinline__ float operator()(float id, float v) const {
    return __builtin_bit_cast(float, dpp_i32<CTRL, ROWM, BANKM>(__builtin_bit_cast(int, id), __builtin_bit_cast(int, v)));
  }
  __device__ __forceinline__ double operator()(double id, double v) const {
    const long long u = __builtin_bit_cast(long long, v), e = __builtin_bit_cast(long long, id);
    const int lo = dpp_i32<CTRL, ROWM, BANKM>((int)(e & 0xffffffffll), (int)(u & 0xffffffffll));
    const int hi = dpp_i32<CTRL, ROWM, BANKM>((int)(e >> 32), (int)(u >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
  }
};
// lane ^ 16 within each half-wave (ds_swizzle, bit mode)
struct Swz16Move {
  __device__ __forceinline__ int operator()(int, int v) const { return __builtin_amdgcn_ds_swizzle(v, (16 << 10) | 0x1F); }
  __device__ __forceinline__ float operator()(float id, float v) const {
    return __builtin_bit_cast(float, (*this)(0, __builtin_bit_cast(int, v)));
  }
  __device__ __forceinline__ double operator()(double, double v) const {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = (*this)(0, (int)(u & 0xffffffffll));
    const int hi = (*this)(0, (int)(u >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
  }
};
__device__ __forceinline__ int bcast63(int v) { return __builtin_amdgcn_readlane(v, 63); }
__device__ __forceinline__ float bcast63(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}
__device__ __forceinline__ double bcast63(double v) {
  const long long u = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), 63);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), 63);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Reductions over the lanes of one IA.  Groups of <= 32 lanes: butterflies
// (quad permutations, half-row and row mirrors, a half-wave swizzle) that leave
// the total in every lane.  Groups of >= 64 lanes: inclusive prefix steps
// within each 16-lane row (row_shr 1, 2, 4, 8), then the row broadcasts of
// lanes 15 and 31, which leave the wave's total in lane 63; one wave is then
// broadcast from lane 63, several are combined through LDS in wave order.
// Every lane ends with the same value, from the same order of operations.
template <int TPI>
struct Group {
  static_assert(TPI == 1 || TPI == 2 || TPI == 4 || TPI == 8 || TPI == 16 || TPI == 32 || TPI % 64 == 0,
                "a power of two up to 32 lanes, or whole waves");
  static constexpr int NW = (TPI + 63) / 64;

  // One reduction over the group: f(mv) is called once per step, with mv(id, v)
  // the step's partner value.
  template <typename F>
  __device__ __forceinline__ static void steps(F&& f) {
    if constexpr (TPI <= 32) {
      if constexpr (TPI > 1) f(DppMove<0xB1, 0xF, 0xF>{});   // quad_perm [1,0,3,2]
      if constexpr (TPI > 2) f(DppMove<0x4E, 0xF, 0xF>{});   // quad_perm [2,3,0,1]
      if constexpr (TPI > 4) f(DppMove<0x141, 0xF, 0xF>{});  // row_half_mirror
      if constexpr (TPI > 8) f(DppMove<0x140, 0xF, 0xF>{});  // row_mirror
      if constexpr (TPI > 16) f(Swz16Move{});
    } else {
      f(DppMove<0x111, 0xF, 0xF>{});  // row_shr:1
      f(DppMove<0x112, 0xF, 0xF>{});  // row_shr:2
      f(DppMove<0x114, 0xF, 0xF>{});  // row_shr:4
      f(DppMove<0x118, 0xF, 0xF>{});  // row_shr:8
      f(DppMove<0x142, 0xA, 0xF>{});  // row_bcast:15 into rows 1 and 3
      f(DppMove<0x143, 0xC, 0xF>{});  // row_bcast:31 into rows 2 and 3
    }
  }

  template <typename V, typename Op>
  __device__ __forceinline__ static V reduce(V v, Op op, double* red, int t) {
    steps([&](auto mv) { v = op(v, mv((V)Op::id(), v)); });
    if constexpr (TPI == 64) {
      v = bcast63(v);
    } else if constexpr (TPI > 64) {
      V* rv = reinterpret_cast<V*>(red);
      __syncthreads();
      if ((t & 63) == 63) rv[t >> 6] = v;
      __syncthreads();
      V r = rv[0];
#pragma unroll 4
      for (int w = 1; w < NW; ++w) r = op(r, rv[w]);
      v = r;
    }
    return v;
  }

  // Up to three reductions sharing one exchange through LDS (slots 0-15,
  // 16-31, 32-47); N = 2 leaves c untouched.
  template <int N, typename OpA, typename OpB, typename OpC>
  __device__ __forceinline__ static void reduce_n(double& a, double& b, double& c, OpA opa, OpB opb, OpC opc,
                                                  double* red, int t) {
    steps([&](auto mv) {
      a = opa(a, mv(OpA::id(), a));
      b = opb(b, mv(OpB::id(), b));
      if constexpr (N > 2) c = opc(c, mv(OpC::id(), c));
    });
    if constexpr (TPI == 64) {
      a = bcast63(a);
      b = bcast63(b);
      if constexpr (N > 2) c = bcast63(c);
    } else if constexpr (TPI > 64) {
      __syncthreads();
      if ((t & 63) == 63) {
        red[t >> 6] = a;
        red[16 + (t >> 6)] = b;
        if constexpr (N > 2) red[32 + (t >> 6)] = c;
      }
      __syncthreads();
      double ra = red[0], rb = red[16], rc = N > 2 ? red[32] : 0.0;
#pragma unroll 4
      for (int w = 1; w < NW; ++w) {
        ra = opa(ra, red[w]);
        rb = opb(rb, red[16 + w]);
        if constexpr (N > 2) rc = opc(rc, red[32 + w]);
      }
      a = ra;
      b = rb;
      if constexpr (N > 2) c = rc;
    }
  }
  template <typename OpA, typename OpB>
  __device__ __forceinline__ static void reduce2(double& a, double& b, OpA opa, OpB opb, double* red, int t) {
    double unused = 0.0;
    reduce_n<2>(a, b, unused, opa, opb, opb, red, t);
  }
};

struct OpAdd {
  __device__ static constexpr double id() { return 0.0; }
  __device__ double operator()(double a, double b) const { return a + b; }
};
struct OpMax {
  __device__ static constexpr double id() { return -__builtin_inf(); }
  __device__ double operator()(double a, double b) const { return a > b ? a : b; }
};
// max that propagates NaN (Julia's `maximum`)
struct OpMaxNaN {
  __device__ static constexpr double id() { return -__builtin_inf(); }
  __device__ double operator()(double a, double b) const {
    return (a != a || b != b) ? (a + b) : (a > b ? a : b);
  }
};

// (value, linear index) argmax with the reference's tie-break: the first element
// in column-major order whose value is strictly above everything before it,
// starting from 0 (common_ops.jl:112-121).  i = INT32_MAX means "none".
template <typename T>
struct Best {
  T v;
  int32_t i;
};
template <typename T>
__device__ __forceinline__ Best<T> best_of(Best<T> a, Best<T> b) {
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

// Offsets of one IA in one dim (0-based).
// Search-region shift of IA ia (index within its job) along dim k, whose
// vector-field coordinate is v: 0 without a shift array, else clamped so that the
// shifted IA [TLF + d, BRB + d] stays inside the image.
__device__ __forceinline__ int32_t shift_of(const Job& J, int32_t ia, int k, int32_t v) {
  if (!J.shift) return 0;
  const int32_t d = J.shift[3 * ia + k];
  const int32_t tlf = v * J.st[k];
  const int32_t lo = -tlf, hi = (int32_t)J.dims[k] - J.is[k] - tlf;
  return d < lo ? lo : (d > hi ? hi : d);
}

struct Geo {
  int64_t tlf;     // IA start
  int64_t sr;      // clipped search-region start
  int32_t gsz;     // clipped search-region length
  int32_t goff;    // placement of the search region in the padded buffer (piv_ops.jl:72)
  int32_t off_hi;  // SR_BRB_off
};

// In 32-bit arithmetic: coordinates stay below dim < 2^31 (the host checks
// the sizes), and the scalar unit that evaluates this for every wave of a
// workgroup is shared by the CU's SIMDs.
// d shifts the search region (multipass): it is the region the reference builds
// for an IA at TLF + d, while the IA itself stays at TLF.  The caller keeps the
// shifted IA inside the image (shift_of).
__device__ __forceinline__ Geo geo_of(int32_t v, int32_t is, int32_t sm, int32_t st, int32_t dim,
                                      int32_t d = 0) {
  Geo g;
  const int32_t tlf = v * st;     // TLF - 1
  const int32_t stl = tlf + d;    // TLF - 1 of the search region's IA
  const int32_t brb = stl + is - 1;  // BRB - 1
  const int32_t lo = stl - sm;
  const int32_t hi = brb + sm;
  const int32_t sr = lo > 0 ? lo : 0;
  const int32_t srb = hi < dim - 1 ? hi : dim - 1;
  g.tlf = tlf;
  g.sr = sr;
  g.gsz = srb - sr + 1;
  // SR_TLF_off = |min(0, TLF - (SM - 1))| with 1-based TLF (piv_ops.jl:72)
  const int32_t a = (stl + 1) - (sm - 1);
  g.goff = a < 0 ? -a : 0;
  // SR_BRB_off = |min(0, dim - (BRB + SM))| (piv_ops.jl:73)
  const int32_t b = dim - ((brb + 1) + sm);
  g.off_hi = b < 0 ? -b : 0;
  return g;
}

// Which job of a batch owns global IA index ia.
__device__ __forceinline__ int find_job(const int64_t* __restrict__ prefix, int n_jobs, int64_t ia) {
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= ia) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// XCD-aware remap: consecutive hardware block ids land on different XCDs; give
// each XCD a contiguous range of IAs so neighbouring windows share its L2.
__device__ __forceinline__ int32_t xcd_remap32(int32_t b, int32_t nb) {
  const int32_t q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return x * q + (x < r ? x : r) + k;
}

// LDS layout of one IA: the L x L complex transform buffer (which after the
// transform holds the similarity window), then the first correlation of the
// masked measure, then the reduction slots.
__host__ __device__ inline size_t sim_bytes(int ns0, int ns1, int tsize) {
  return (((size_t)ns0 * ns1 * tsize + 15) & ~(size_t)15);
}
__host__ __device__ inline size_t region_bytes_2d(int L, int tsize, int ns0, int ns1, int ss0) {
  (void)ns0;
  (void)ns1;
  (void)ss0;
  return ((size_t)L * (L + 1) * 2 * tsize + 15) & ~(size_t)15;
}
constexpr int kRedBytes = 512;  // reduction slots per IA (4 x 16 doubles)

// FP32 masked NSQECC with a refine list: an IA goes to the FP64 transforms when
// its values could move an output past the FP64 results' bar.  fr(s) is the
// bound on the normalised value's error per unit transform error (above); over
// ~7,000 IAs of four masked particle pairs (interSize 16 to 64, 20 ellipses
// covering 30%) the peak values moved by at most 4.5e-8 fr, and the SN
// window's mean by at most 2.7e-7 mean(fr) (relative to the mean):
// * peak: flagged when the runner-up is within kMaskRefinePeak fr(peak) of it;
// * SN: flagged when kMaskRefineEps mean(fr) exceeds kMaskRefineSn |mean|;
// * fit: flagged when a log error of 10 kMaskRefineEps fr(peak) could move the
//   subpixel offset by 1e-3 px.
constexpr double kMaskRefinePeak = 1e-6;
constexpr double kMaskRefineEps = 3e-7;
constexpr double kMaskRefineSn = 1e-5;

// ---------------------------------------------------------------------------
// The correlation of one IA.  LPL lanes (a "group", q = lane within it) share a
// line of L samples, M = L/LPL each.  On entry lane (line = k0, q) holds the
// packed window Z = A + iB at k1 = LPL*k + q, k < M.  The forward transform runs
// along dim 2 in registers, then along dim 1 through an LDS transpose; the
// product P = conj(A^) B^ reads each partner bin back from LDS; the inverse runs
// along dim 1 (keeping s0 < ns0) and dim 2 (keeping s1 < ns1).  On exit the
// fully-overlapping window of real correlation values (scaled by 1/L^2) is at
// the start of the Z region as T, column-major with ns0 fastest.
template <int L, int LPL, typename T>
__device__ __forceinline__ void line_fwd(cpx<T> (&x)[L / LPL], int q) {
  if constexpr (LPL == 1) fft_line<L, -1>(x);
  else fft_group_dit<L, LPL, -1, T>(x, q, DppXchg{});
}
template <int L, int LPL, typename T>
__device__ __forceinline__ void line_inv(cpx<T> (&x)[L / LPL], int q) {
  if constexpr (LPL == 1) fft_line<L, 1>(x);
  else fft_group_dif<L, LPL, 1, T>(x, q, DppXchg{});
}

// Lanes per line for the last inverse pass (rows along dim 2, two per
// transform): twice the forward passes' when the launch has the lanes for the
// most row pairs a length-L transform can need (ns0 <= L/2 + 1) and the
// per-lane length has a plan.  That pass has few lines, so shorter per-lane
// chains shorten it.
__host__ __device__ constexpr bool has_plan(int m) {
  return m == 4 || m == 6 || m == 8 || m == 12 || m == 16 || m == 24 || m == 32 || m == 48 || m == 64 ||
         m == 96 || m == 128;
}
template <int L, int LPL, int TPI>
__host__ __device__ constexpr int c2r_lanes() {
  constexpr int P = 2 * LPL;
  return (P <= 16 && P * ((L / 2 + 2) / 2) <= TPI && has_plan(L / P)) ? P : LPL;
}
// A wider group for the same pass, used when an IA's line pairs fit the block
// (checked at run time); c2r_lanes itself when there is none.
template <int L, int LPL, int TPI>
__host__ __device__ constexpr int c2r_wide_lanes() {
  constexpr int P = 2 * c2r_lanes<L, LPL, TPI>();
  return (P <= 16 && P <= TPI && has_plan(L / P)) ? P : c2r_lanes<L, LPL, TPI>();
}

template <int L, int LPL, typename T, typename ZB>
__device__ __forceinline__ void correlate_2d(cpx<T> (&x)[L / LPL], const ZB& Z, int t, int ns0, int ns1) {
  constexpr int LD = L + 1;
  constexpr int M = L / LPL;
  const int line = t / LPL;
  const int q = t % LPL;
  const int cM = (LPL == 1 ? 0 : chunk_of<LPL>(q)) * M;  // this lane's chunk after a forward pass
  const bool lane = line < L;

  // --- forward along dim 2 (registers): line = k0
  if (lane) {
    line_fwd<L, LPL, T>(x, q);
#pragma unroll
    for (int k = 0; k < M; ++k) Z.st(cM + k, line, x[k]);
  }
  __syncthreads();
  // --- forward along dim 1: line = k1
  if (lane) {
#pragma unroll
    for (int k = 0; k < M; ++k) x[k] = Z.ld(line, LPL * k + q);
    line_fwd<L, LPL, T>(x, q);
#pragma unroll
    for (int k = 0; k < M; ++k) Z.st(line, cM + k, x[k]);
  }
  __syncthreads();
  // The correlation is real, so P is Hermitian: the dim-1 inverse of column
  // L - k1 is the conjugate of column k1's.  Only columns k1 <= L/2 are
  // multiplied and inverted; the dim-2 pass then rebuilds each row's full
  // spectrum from that half and inverts two real rows per complex transform.
  constexpr int KH = L / 2 + 1;  // columns kept
  const bool half = line < KH;
  if (half) {
    // --- P(k) = conj(A^(k)) B^(k) with Z^ = A^ + i B^ (fft_ops.jl:95-105)
    const int k1n = (L - line) % L;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int k0 = cM + k;
      const int k0n = (L - k0) % L;
      cpx<T> z = x[k];
      cpx<T> n = Z.ld(k1n, k0n);
      T ar = T(0.5) * (z.x + n.x), ai = T(0.5) * (z.y - n.y);
      T br = T(0.5) * (z.y + n.y), bi = T(0.5) * (n.x - z.x);
      x[k] = cpx<T>{ar * br + ai * bi, ar * bi - ai * br};
    }
    // --- inverse along dim 1: lane q then holds s0 = LPL*m + q
    line_inv<L, LPL, T>(x, q);
  }
  __syncthreads();
  if (half) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int s0 = LPL * m + q;
      if (s0 < ns0) Z.st(line, s0, x[m]);
    }
  }
  __syncthreads();
  // --- inverse along dim 2, rows s0 = 2r and 2r + 1 packed as Ya + i Yb:
  // line = r, lane q2 holds k1 = cM2 + k of the packed spectrum (P2 lanes a line)
  const int npair = (ns0 + 1) / 2;
  auto c2r = [&](auto p2c) {
    constexpr int P2 = decltype(p2c)::value;
    constexpr int M2 = L / P2;
    const int line2 = t / P2, q2 = t % P2;
    const int cM2 = (P2 == 1 ? 0 : chunk_of<P2>(q2)) * M2;
    const bool act = line2 < npair;
    cpx<T> y[M2];
    if (act) {
      const int sa = 2 * line2, sb = 2 * line2 + 1;
      const bool has_b = sb < ns0;
#pragma unroll
      for (int k = 0; k < M2; ++k) {
        const int k1 = cM2 + k;
        const bool mirror = k1 >= KH;
        const int kc = mirror ? L - k1 : k1;
        const cpx<T> ya = Z.ld(kc, sa);
        const cpx<T> yb = has_b ? Z.ld(kc, sb) : cpx<T>{T(0), T(0)};
        // Y(s, k1) = conj(Y(s, L - k1)) above L/2
        const T ayi = mirror ? -ya.y : ya.y, byi = mirror ? -yb.y : yb.y;
        y[k] = cpx<T>{ya.x - byi, ayi + yb.x};
      }
      line_inv<L, P2, T>(y, q2);
    }
    __syncthreads();
    if (act) {
      const auto S = Z.real();
      const T kScale = T(1) / T(L * L);
      const int sa = 2 * line2, sb = 2 * line2 + 1;
#pragma unroll
      for (int m = 0; m < M2; ++m) {
        const int s1 = P2 * m + q2;
        if (s1 < ns1) {
          S[s1 * ns0 + sa] = y[m].x * kScale;
          if (sb < ns0) S[s1 * ns0 + sb] = y[m].y * kScale;
        }
      }
    }
    __syncthreads();
  };
  // P2 lanes per line fit every margin this length allows; when the block holds
  // one IA (so the choice is block-uniform) and this IA's line pairs fit, twice
  // as many lanes halve each lane's serial chain
  constexpr int TPI = Cfg2<L, T, LPL>::TPI;
  constexpr int P2 = c2r_lanes<L, LPL, TPI>();
  constexpr int P2w = c2r_wide_lanes<L, LPL, TPI>();
  if constexpr (P2w != P2 && Cfg2<L, T, LPL>::IPB == 1) {
    if (npair * P2w <= TPI) c2r(std::integral_constant<int, P2w>{});
    else c2r(std::integral_constant<int, P2>{});
  } else {
    c2r(std::integral_constant<int, P2>{});
  }
}

// Box sum of an FP64 integral array over image rows [a0, b0) x [a1, b1).
__device__ __forceinline__ double sat_box(const double* __restrict__ S, int64_t D0, int64_t a0, int64_t b0,
                                          int64_t a1, int64_t b1) {
  return ((S[b0 + D0 * b1] - S[a0 + D0 * b1]) - S[b0 + D0 * a1]) + S[a0 + D0 * a1];
}

// Waves per SIMD that the LDS footprint of one block allows (at most 8): the
// occupancy the register allocation is asked to fit, so that VGPRs do not
// become the tighter limit.  The masked measure's extra correlation window is
// sized for the largest margin a length-L transform can hold (ns <= L/2 + 1).
#ifndef MQPIV_OCC_CAP
#define MQPIV_OCC_CAP 8
#endif
template <int L, int ALG, typename T, int LPL>
constexpr int min_waves_2d() {
  // global transform buffer: one IA's block per CU, whose waves all fit the
  // SIMDs together (16 waves: 4 per SIMD, 128 VGPRs)
  if constexpr (gz_len<L, T>()) return Cfg2<L, T, LPL>::NT / 256 > 0 ? Cfg2<L, T, LPL>::NT / 256 : 1;
  using C = Cfg2<L, T, LPL>;
  constexpr size_t z = ((size_t)L * (L + 1) * 2 * sizeof(T) + 15) & ~(size_t)15;
  constexpr size_t c1 = ALG == ALG_MASK_NSQECC ? (((size_t)(L / 2 + 1) * (L / 2 + 1) * sizeof(T) + 15) & ~(size_t)15) : 0;
  constexpr size_t per_block = (z + c1 + kRedBytes) * C::IPB;
  constexpr int blocks = (int)((160 * 1024) / per_block);
  constexpr int waves = blocks * (C::NT / 64) / 4;
  // A lower target only pays in whole blocks per CU: FP32 ZNCC at length 96
  // (12-wave blocks, 6 waves per SIMD by LDS) spills 2 VGPRs at 80, but at 5
  // waves per SIMD only one block fits a CU and the kernel runs 1.6x longer;
  // masked FP32 at lengths 32/64 (4-wave blocks) spills a little at 128 VGPRs
  // and still runs 10% faster than at 3 waves per SIMD.
  return waves < 1 ? 1 : (waves > MQPIV_OCC_CAP ? MQPIV_OCC_CAP : waves);
}

// One IA per block-slot.  SH: search regions shifted per IA (Job::shift,
// multipass); a separate instantiation, so that kernels without shifts carry
// none of its registers.
// gz / gz_slots: the global parts of the transform buffers of a gz_len launch
// (gz_slots of them after the slot flags), else unused.
// LIST: the IAs are the entries of a device list (refine_list layout: count at
// list[0], global IA indices from list[kRefineHead]), written by an earlier
// kernel on the stream; the grid is persistent and every block steps through
// the list by the grid size until the count is reached.
// refine_out (FP32 masked NSQECC): a refine list the kernel appends the IAs to
// whose values may miss the FP64 results' bar, or nullptr.
template <int L, int ALG, typename T, int LPL, bool SH, bool LIST>
__global__ __launch_bounds__((Cfg2<L, T, LPL>::NT), (min_waves_2d<L, ALG, T, LPL>())) void piv2d_kernel(Job j0, const Job* __restrict__ jobs,
                                                                      const int64_t* __restrict__ prefix,
                                                                      int n_jobs, int64_t n_ia_total, int ns0_cap,
                                                                      int ns1_cap, char* gz, int gz_slots,
                                                                      const int32_t* __restrict__ list,
                                                                      int32_t* __restrict__ refine_out) {
  const int32_t vb = xcd_remap32((int32_t)blockIdx.x, (int32_t)gridDim.x);
  using C = Cfg2<L, T, LPL>;
  constexpr int H = C::H;
  constexpr bool GZ = gz_len<L, T>();
  static_assert(!GZ || C::IPB == 1, "a global transform buffer holds one IA");
  static_assert(!(LIST && SH), "list launches take unshifted regions");
  // FP32 masked NSQECC with a refine list: the per-IA quantities that bound its
  // values' error (below) are gathered as well
  constexpr bool kMaskRefine = ALG == ALG_MASK_NSQECC && std::is_same<T, float>::value && !LIST;
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t lpos = vb;  // LIST: this block's current list position

  // without LIST the body runs once (a loop keeps the register allocation of
  // this body as measured)
  for (;;) {
  int tid = (int)threadIdx.x;
  // LIST: the thread index is re-read every iteration, so that values derived
  // from it are not hoisted out of the loop and held across the body
  if constexpr (LIST) __asm__ volatile("" : "+v"(tid));
  // one IA per block: every IA quantity below is uniform over the block, which
  // lets the compiler keep it in scalar registers
  const int t = C::IPB == 1 ? tid : tid % C::TPI;
  const int sub = C::IPB == 1 ? 0 : tid / C::TPI;
  const int line = t / LPL;
  const int h = t % LPL;  // lane within the line's group

  // IA indices in 32-bit arithmetic (the host keeps a launch below 2^31 IAs)
  const int32_t nt = LIST ? list[0] : (int32_t)n_ia_total;
  if (LIST && lpos * C::IPB >= nt) break;  // uniform over the block
  const int32_t ia_g = (LIST ? lpos : vb) * C::IPB + sub;
  const bool valid = ia_g < nt;
  const int32_t ia_c = LIST ? list[kRefineHead + (valid ? ia_g : nt - 1)] : (valid ? ia_g : nt - 1);
  Job J = j0;
  int32_t ia32 = ia_c;
  if (jobs) {  // batched launch: locate the owning job
    const int jb = find_job(prefix, n_jobs, ia_c);
    J = jobs[jb];
    ia32 = ia_c - (int32_t)prefix[jb];
  }
  const int64_t ia = ia32;

  const int sm0 = J.sm[0], sm1 = J.sm[1];
  const int is0 = J.is[0], is1 = J.is[1];
  const int ns0 = 2 * sm0 + 1, ns1 = 2 * sm1 + 1;
  const int ss0 = is0 + 2 * sm0;

  // LDS layout is uniform across the launch (jobs of a batch may differ in
  // their margins): the host passes the largest ns0, ns1 of the launch
  constexpr int LDc = L + 1;
  const size_t zbytes = GZ ? 0 : region_bytes_2d(L, sizeof(T), ns0_cap, ns1_cap, ss0);
  const size_t c1bytes = (ALG == ALG_MASK_NSQECC) ? sim_bytes(ns0_cap, ns1_cap, sizeof(T)) : 0;
  const size_t per_ia = zbytes + c1bytes;
  unsigned char* base = smem + (size_t)sub * per_ia;
  cpx<T>* Z = reinterpret_cast<cpx<T>*>(base);
  using ZB = std::conditional_t<GZ, ZGlob<T, LDc>, ZLds<T, LDc>>;
  ZB zb;
  if constexpr (!GZ) zb.z = Z;
  T* C1 = reinterpret_cast<T*>(base + zbytes);
  double* red = reinterpret_cast<double*>(smem + (size_t)C::IPB * per_ia + (size_t)sub * kRedBytes);
  // Global part of the buffer: a slot claimed for this block.  The slot flags
  // (0 free, 1 held) precede the buffers; the host provides twice as many slots
  // as blocks can be resident, so a free one always exists and the probe
  // (bounded all the same) ends at once in practice.  The slot is released
  // after the block's last access (end of this function).
  int32_t gz_slot = 0;
  if constexpr (GZ) {
    // the claim is passed on through the first reduction slot (every thread
    // reads it before a reduction writes the slots again, behind a barrier)
    int32_t& claimed = *reinterpret_cast<int32_t*>(red);
    int32_t* flags = reinterpret_cast<int32_t*>(gz);
    if (tid == 0) {
      int32_t got = -1;
      const uint32_t n = (uint32_t)gz_slots;
      for (uint32_t i = 0; i < kGzProbeLimit && got < 0; ++i) {
        const uint32_t s = ((uint32_t)blockIdx.x + i) % n;
        if (atomicCAS(&flags[s], 0, 1) == 0) got = (int32_t)s;
      }
      __threadfence();
      claimed = got;
    }
    __syncthreads();
    gz_slot = claimed;
    if (gz_slot < 0) {  // no slot (cannot happen with the host's slot count): NaN outputs
      if (valid && t == 0) {
        if (J.out_dtype == DT_F32) {
          ((float*)J.vf)[2 * ia32] = NAN;
          ((float*)J.vf)[2 * ia32 + 1] = NAN;
          if (J.sn) ((float*)J.sn)[ia32] = NAN;
        } else {
          ((double*)J.vf)[2 * ia32] = NAN;
          ((double*)J.vf)[2 * ia32 + 1] = NAN;
          if (J.sn) ((double*)J.sn)[ia32] = NAN;
        }
      }
      return;
    }
    zb.g = (gmut<cpx<T>>)(gz + kGzFlagBytes + (size_t)gz_slot * gz_stride_2d<L, T>());
  }
  const auto S = zb.real();

  const uint32_t vn0 = (uint32_t)J.vf_n[0];
  const int32_t v1 = (int32_t)((uint32_t)ia32 / vn0);
  const int32_t v0 = ia32 - v1 * (int32_t)vn0;
  const int32_t sh0 = SH ? shift_of(J, ia32, 0, v0) : 0, sh1 = SH ? shift_of(J, ia32, 1, v1) : 0;
  const Geo g0 = geo_of(v0, is0, sm0, J.st[0], (int32_t)J.dims[0], sh0);
  const Geo g1 = geo_of(v1, is1, sm1, J.st[1], (int32_t)J.dims[1], sh1);
  const int dt = J.in_dtype;
  const int64_t d0 = J.dims[0];

  // Lane (line = k0, q) covers k1 = LPL*k + q, k < H.  Every lane issues all its
  // loads before using any (one wait for the batch), and every load reads an
  // in-image pixel; f_at / g_at discard the values outside the IA / the clipped
  // search region.  Two address forms (the host keeps (dims[0] + 2) * 1024 < 2^32,
  // so both fit 32-bit offsets):
  // * FP64 transforms: a block-uniform base (the IA's first pixel; the clipped
  //   search region's first pixel) plus a per-lane byte offset; rows and columns
  //   outside the window are clamped to its nearest row / column, and column
  //   k1 = LPL*k + h lies k * LPL * d0 elements after column h (an add and a
  //   clamp per load instead of a multiply, selects and a 64-bit add);
  // * FP32 transforms: a per-lane base plus d0 * k1, with positions outside the
  //   window sent to one pixel (the IA origin, or the search region's first row /
  //   column), which measured faster there than the clamped form.
  constexpr bool kClampedGather = std::is_same<T, double>::value;
  const int k0 = line;
  const bool f_row = k0 < is0;
  const bool g_row = k0 >= g0.goff && k0 < g0.goff + g0.gsz;
  auto k1_of = [&](int k) { return LPL * k + h; };
  auto f_at = [&](int k1) -> bool { return f_row && k1 < is1; };
  auto g_at = [&](int k1) -> bool { return g_row && k1 >= g1.goff && k1 < g1.goff + g1.gsz; };
  const int64_t f_org = g0.tlf + d0 * g1.tlf;
  const int64_t g_org = g0.sr + d0 * g1.sr;
  const int32_t id0 = (int32_t)d0;
  const int32_t f_r = k0 < is0 - 1 ? k0 : is0 - 1;
  const int32_t g_r = min(max(k0 - g0.goff, 0), g0.gsz - 1);
  const int32_t f_c0 = id0 * h, f_cmax = id0 * (is1 - 1);
  const int32_t g_c0 = id0 * (h - g1.goff), g_cmax = id0 * (g1.gsz - 1);
  const int32_t c_step = id0 * LPL;
  auto f_el = [&](int k) -> uint32_t {
    const int32_t c = f_c0 + k * c_step;
    return (uint32_t)(f_r + (c < f_cmax ? c : f_cmax));
  };
  auto g_el = [&](int k) -> uint32_t { return (uint32_t)(g_r + min(max(g_c0 + k * c_step, 0), g_cmax)); };
  auto at_byte = [](const void* base, int64_t org, uint32_t el, auto tin) {
    using TIN = decltype(tin);
    const gptr<char> b = as_global<char>(base) + org * (int64_t)sizeof(TIN);
    return *(gptr<TIN>)(b + el * (uint32_t)sizeof(TIN));
  };
  const uint32_t ud0 = (uint32_t)d0;
  const int64_t f_lane = f_row ? (g0.tlf + k0) + d0 * g1.tlf : f_org;
  const int64_t g_lane = g_row ? (g0.sr + (k0 - g0.goff)) + d0 * (g1.sr - g1.goff) : g0.sr + d0 * (g1.sr - g1.goff);
  auto f_off = [&](int k1) -> uint32_t { return f_at(k1) ? ud0 * (uint32_t)k1 : 0u; };
  auto g_off = [&](int k1) -> uint32_t { return ud0 * (uint32_t)(g_at(k1) ? k1 : g1.goff); };
  // F (or mask) and G sample of column k1_of(k) for this lane
  auto ld_f = [&](const void* base, int k, auto tin) {
    using TIN = decltype(tin);
    if constexpr (kClampedGather) return at_byte(base, f_org, f_el(k), tin);
    else return (as_global<TIN>(base) + f_lane)[f_off(k1_of(k))];
  };
  auto ld_g = [&](const void* base, int k, auto tin) {
    using TIN = decltype(tin);
    if constexpr (kClampedGather) return at_byte(base, g_org, g_el(k), tin);
    else return (as_global<TIN>(base) + g_lane)[g_off(k1_of(k))];
  };

  // ZNCC means over the IA and the clipped search region (zncc_ops.jl:10,22),
  // computed in the input element type, as the reference does, from the loaded window
  float meanF32 = 0.f, meanG32 = 0.f;
  double meanF64 = 0.0, meanG64 = 0.0;

  // ---- load the packed window into registers; sum(F^2) and the IA maxima on the way
  cpx<T> x[H];
  double sumF2 = 0.0, mx1 = -INFINITY, mx2 = -INFINITY;
  // FP64 ZNCC forms its window energies and sum(F'^2) from the rounded
  // differences G - mean, F - mean themselves (below), as the reference does
  constexpr bool kLocalW = ALG == ALG_ZNCC && std::is_same<T, double>::value;
  double sa2 = 0.0;
  // kMaskRefine: sums of (F m)^2, G^2, m^2 and G^4 over the IA / search window
  float e_fm = 0.f, e_g2 = 0.f, e_m = 0.f, e_g4 = 0.f;
  // G'^2 of the lane's window samples, kept through the correlation for the
  // table (lengths whose buffer is in LDS; the others read the image again)
  constexpr bool kKeepG2 = kLocalW && !GZ;
  double gq[kKeepG2 ? H : 1];
  const bool want_max = J.has_threshold || J.has_mask_threshold;
  // Values are loaded in batches of LB per lane: all of a lane's columns at
  // once, except with global transform buffers, whose long lines would hold too
  // many loaded values in registers (there ZNCC loads its window twice: once
  // for the means, once for the packed window).
  constexpr int LB = GZ ? 8 : H;
  static_assert(H % LB == 0, "whole batches");
  auto load = [&](auto tin) {
    using TIN = decltype(tin);
    TIN fv[LB], gv[LB], mv[LB];
    auto fetch = [&](int c0) {
#pragma unroll
      for (int j = 0; j < LB; ++j) {
        fv[j] = ld_f(J.img1, c0 + j, tin);
        gv[j] = ld_g(J.img2, c0 + j, tin);
      }
      if constexpr (ALG == ALG_MASK_NSQECC) {
#pragma unroll
        for (int j = 0; j < LB; ++j) mv[j] = ld_f(J.mask, c0 + j, tin);
      }
    };
    if constexpr (ALG == ALG_ZNCC) {
      // sums of F, G and F^2 in one reduction; sum((F - mean)^2) follows from
      // them (exact algebra; it differs from summing the squares of the
      // rounded differences by a rounding-level factor common to the whole IA,
      // which leaves the peak and SN unchanged)
      double sf = 0.0, sg = 0.0, sff = 0.0;
#pragma unroll
      for (int c0 = 0; c0 < H; c0 += LB) {
        fetch(c0);
#pragma unroll
        for (int j = 0; j < LB; ++j) {
          const int k1 = k1_of(c0 + j);
          const double f = f_at(k1) ? (double)fv[j] : 0.0;
          sf += f;
          sff += f * f;
          if (g_at(k1)) sg += (double)gv[j];
        }
      }
      Group<C::TPI>::template reduce_n<3>(sf, sg, sff, OpAdd{}, OpAdd{}, OpAdd{}, red, t);
      const int64_t nF = (int64_t)is0 * is1;
      const int64_t nG = (int64_t)g0.gsz * g1.gsz;
      meanF32 = (float)sf / (float)nF;
      meanG32 = (float)sg / (float)nG;
      meanF64 = sf / (double)nF;
      meanG64 = sg / (double)nG;
      const double mu = std::is_same<TIN, float>::value ? (double)meanF32 : meanF64;
      const double v = (sff - 2.0 * mu * sf) + (double)nF * mu * mu;
      sumF2 = v > 0.0 ? v : 0.0;
    }
#pragma unroll
    for (int c0 = 0; c0 < H; c0 += LB) {
      if (ALG != ALG_ZNCC || LB != H) fetch(c0);  // one batch: ZNCC reuses it
#pragma unroll
      for (int j = 0; j < LB; ++j) {
        const int k = c0 + j;
        const int k1 = k1_of(k);
        double a = 0.0, g = 0.0;
        if (f_at(k1)) {
          const TIN f = fv[j];
          if (want_max) mx1 = fmax(mx1, (double)f);
          if constexpr (ALG == ALG_ZNCC) {
            a = std::is_same<TIN, float>::value ? (double)((float)f - meanF32) : (double)f - meanF64;
            if constexpr (kLocalW) sa2 += a * a;
          } else if constexpr (ALG == ALG_MASK_NSQECC) {
            const double m = (double)mv[j];
            if (want_max) mx2 = fmax(mx2, m);
            a = (double)f * m;  // masked_nsqecc_ops.jl:19-25
            const double wi = m > 0 ? m : 1.0;
            sumF2 += a * a / wi;  // mask_nsqecc.jl:144-149
            if constexpr (kMaskRefine) e_fm += (float)(a * a);
          } else {
            a = (double)f;
            if constexpr (ALG != ALG_FFTCC) sumF2 += a * a;
          }
        }
        if (g_at(k1)) {
          if constexpr (ALG == ALG_ZNCC)
            g = std::is_same<TIN, float>::value ? (double)((float)gv[j] - meanG32) : (double)gv[j] - meanG64;
          else
            g = (double)gv[j];
          if constexpr (kMaskRefine) e_g2 += (float)(g * g);
        }
        if constexpr (kKeepG2) gq[k] = g * g;
        x[k] = cpx<T>{(T)a, (T)g};
      }
    }
  };
  if (dt == DT_F32) load(float{});
  else load(double{});
  if constexpr (ALG == ALG_NSQECC || ALG == ALG_MASK_NSQECC) sumF2 = Group<C::TPI>::reduce(sumF2, OpAdd{}, red, t);
  // FP64 ZNCC: sum(F'^2) of the rounded differences themselves, as zncc.jl:131-135
  if constexpr (kLocalW) sumF2 = Group<C::TPI>::reduce(sa2, OpAdd{}, red, t);

  // ---- skip test (piv_ops.jl:134-157)
  int skip = 0;
  if (J.skip) {
    skip = J.skip[ia] != 0;
  } else if (want_max) {
    // (mx2, the mask maximum, is only gathered by the masked measure, which is
    // the only one the host gives a mask threshold)
    Group<C::TPI>::reduce2(mx1, mx2, OpMax{}, OpMax{}, red, t);
    if (J.has_threshold && mx1 < J.threshold) skip = 1;
    if (J.has_mask_threshold && mx2 < J.mask_threshold) skip = 1;
  }

  // ---- correlation(s)
  correlate_2d<L, LPL, T>(x, zb, t, ns0, ns1);
  if constexpr (ALG == ALG_MASK_NSQECC) {
    for (int i = t; i < ns0 * ns1; i += C::TPI) C1[i] = S[i];
    // second correlation: mask (x) G^2 (mask_nsqecc.jl:155)
    auto load2 = [&](auto tin) {
      using TIN = decltype(tin);
#pragma unroll
      for (int c0 = 0; c0 < H; c0 += LB) {
        TIN mv[LB], gv[LB];
#pragma unroll
        for (int j = 0; j < LB; ++j) {
          mv[j] = ld_f(J.mask, c0 + j, tin);
          gv[j] = ld_g(J.img2, c0 + j, tin);
        }
#pragma unroll
        for (int j = 0; j < LB; ++j) {
          const int k = c0 + j;
          const int k1 = k1_of(k);
          double m = 0.0, g2 = 0.0;
          if (f_at(k1)) m = (double)mv[j];
          if (g_at(k1)) {
            const double g = (double)gv[j];
            g2 = g * g;
          }
          if constexpr (kMaskRefine) {
            e_m += (float)(m * m);
            e_g4 += (float)(g2 * g2);
          }
          x[k] = cpx<T>{(T)m, (T)g2};
        }
      }
    };
    if (dt == DT_F32) load2(float{});
    else load2(double{});
    __syncthreads();
    correlate_2d<L, LPL, T>(x, zb, t, ns0, ns1);
  }

  // ---- normalisation, peak and SN in one pass.  Each lane normalises its
  // window elements e = t + j*TPI, writes them back for the Gaussian fit and
  // folds them into its candidates: the peak over [off_lo, 2SM - off_hi] per
  // dim (common_ops.jl:98-124; first strict maximum in column-major order,
  // starting from 0) and the SN maximum and sum over [off_lo, 2SM-1-off_hi]
  // (piv_ops.jl:336-367).  One combined reduction follows.
  const int lo0 = g0.goff, hi0 = 2 * sm0 - g0.off_hi;
  const int lo1 = g1.goff, hi1 = 2 * sm1 - g1.off_hi;
  const int shi0 = 2 * sm0 - 1 - g0.off_hi, shi1 = 2 * sm1 - 1 - g1.off_hi;
  const bool want_sn = J.compute_sn != 0;
  const int nE = ns0 * ns1;
  Best<T> bst{T(0), INT32_MAX};
  double mx = -INFINITY, smv = 0.0;
  // window position (p0, p1) of this lane's elements e = t + m * TPI, m = 0, 1,
  // ...: one division here, then a step of (TPI % ns0, TPI / ns0) with a carry
  const int st0 = C::TPI % ns0, st1 = C::TPI / ns0;
  int w0 = t % ns0, w1 = t / ns0;
  auto step = [&]() {
    w0 += st0;
    w1 += st1;
    if (w0 >= ns0) {
      w0 -= ns0;
      ++w1;
    }
  };
  auto visit = [&](int e, int p0, int p1, T v) {
    if (p0 >= lo0 && p0 <= hi0 && p1 >= lo1 && p1 <= hi1 && v > bst.v) bst = Best<T>{v, e};
    if (want_sn && p0 >= lo0 && p0 <= shi0 && p1 >= lo1 && p1 <= shi1) {
      const double dv = (double)v;
      mx = (dv > mx || dv != dv) ? dv : mx;
      smv += dv;
    }
  };
  if constexpr (kLocalW) {
    // W(s) = sum of G'^2 over the search-buffer window [s, s + is), with G' the
    // padded search window minus its mean, rounded in the input element type
    // (zncc_ops.jl:22-24), from a summed-area table of G'^2 built here in LDS
    // (integralArraySQ!, common_ops.jl:142-180; window sums as common_ops.jl:
    // 203-245).  Forming it from image-wide integral arrays of G and G^2
    // instead changes each window by the rounding of G - mean, which the near
    // cancellation of ZNCC's SN sum can amplify past its bar.
    const int ss1 = is1 + 2 * sm1;
    const int Dl = ss0 + 1;
    double* satl;
    if constexpr (GZ) satl = reinterpret_cast<double*>(smem + (size_t)C::IPB * per_ia + (size_t)C::IPB * kRedBytes);
    else satl = reinterpret_cast<double*>(base + sim_bytes(ns0_cap, ns1_cap, sizeof(T)));
    // G'^2 into the table (entry (i, j) = pixel (i - 1, j - 1) of the padded
    // window; border entries, index 0 in either dim, are zero): consecutive
    // lanes take consecutive pixels down a column, each lane all its pixels at
    // once.
    for (int i = t; i <= ss0; i += C::TPI) satl[i] = 0.0;         // column 0
    for (int j = t + 1; j <= ss1; j += C::TPI) satl[Dl * j] = 0.0; // row 0
    if constexpr (kKeepG2) {
      // lane (line k0, h) holds column k1 = LPL*k + h of row k0
      if (line < ss0) {
#pragma unroll
        for (int k = 0; k < H; ++k) {
          const int k1 = LPL * k + h;
          if (k1 < ss1) satl[(line + 1) + Dl * (k1 + 1)] = gq[k];
        }
      }
    } else {
      constexpr int KF = (L * L + C::TPI - 1) / C::TPI;  // pixels a lane takes, at most
      const int npx = ss0 * ss1;
      const int st0 = C::TPI % ss0, st1 = C::TPI / ss0;
      int i0 = t % ss0, j0 = t / ss0;
      static_assert(KF <= 32, "one data bit per pixel");
      double v[KF];
      int dst[KF];
      uint32_t data = 0;  // bit k: pixel k holds image data
      // every load reads an in-image pixel (positions outside the data are
      // clamped to it and masked afterwards), so that no load is predicated
      // and a lane's loads all go out before it waits
      const int32_t id0 = (int32_t)d0;
      auto fill = [&](auto tin) {
        using TIN = decltype(tin);
        const gptr<TIN> org = as_global<TIN>(J.img2) + (g0.sr + d0 * g1.sr);
        TIN gv[KF];
#pragma unroll
        for (int k = 0; k < KF; ++k) {
          const int e = t + k * C::TPI;
          const int x0 = i0 - g0.goff, x1 = j0 - g1.goff;
          const bool in = e < npx && x0 >= 0 && x0 < g0.gsz && x1 >= 0 && x1 < g1.gsz;
          data |= (in ? 1u : 0u) << k;
          dst[k] = e < npx ? (i0 + 1) + Dl * (j0 + 1) : Dl * (ss1 + 1);  // past the table: a spare entry
          const int32_t cx0 = min(max(x0, 0), g0.gsz - 1), cx1 = min(max(x1, 0), g1.gsz - 1);
          gv[k] = org[(uint32_t)(cx0 + id0 * cx1)];
          i0 += st0;
          j0 += st1;
          if (i0 >= ss0) {
            i0 -= ss0;
            ++j0;
          }
        }
#pragma unroll
        for (int k = 0; k < KF; ++k) {
          const double gg = std::is_same<TIN, float>::value ? (double)((float)gv[k] - meanG32)
                                                            : (double)gv[k] - meanG64;
          v[k] = ((data >> k) & 1u) ? gg * gg : 0.0;  // outside the data: 0, not -mean (zncc_ops.jl:22-24)
        }
      };
      if (dt == DT_F32) fill(float{});
      else fill(double{});
#pragma unroll
      for (int k = 0; k < KF; ++k) satl[dst[k]] = v[k];
    }
    __syncthreads();
    // Prefix sums down the columns, then along the rows: lines of at most L
    // entries split into PS segments of RS entries; lane t scans segment
    // t % PS of line t / PS in registers and adds the totals of the line's
    // earlier segments, taken from the lanes beside it.  A lane writes back
    // only the entries it read, so each pass needs no barrier inside.
    constexpr int PS = C::TPI / L > 1 ? (C::TPI / L > 8 ? 8 : C::TPI / L) : 1;
    constexpr int RS = (L + PS - 1) / PS;
    const int seg = t % PS, ln = t / PS;
    auto seg_offset = [&](double total) {
      double inc = total;
#pragma unroll
      for (int d = 1; d < PS; d <<= 1) {
        const double o = __shfl_up(inc, d, PS);
        if (seg >= d) inc += o;
      }
      return inc - total;
    };
    auto pass = [&](int lines, int len, int line_stride, int elem_stride) {
      for (int lb = 0; lb < lines; lb += C::TPI / PS) {
        const int l = lb + ln;
        // entries past the line (or of lines past the last) read and write
        // the spare entry after the table; their values are never used
        const int spare = Dl * (ss1 + 1);
        const int lbase = (l + 1) * line_stride;
        double w[RS];
        int at[RS];
#pragma unroll
        for (int k = 0; k < RS; ++k) {
          const int m = seg * RS + k;
          at[k] = (l < lines && m < len) ? lbase + (m + 1) * elem_stride : spare;
          w[k] = satl[at[k]];
        }
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < RS; ++k) {
          const int m = seg * RS + k;
          acc += m < len ? w[k] : 0.0;
          w[k] = acc;
        }
        const double off = seg_offset(acc);
#pragma unroll
        for (int k = 0; k < RS; ++k) satl[at[k]] = w[k] + off;
      }
      __syncthreads();
    };
    pass(ss1, ss0, Dl, 1);  // down each column (dim 1)
    pass(ss0, ss1, 1, Dl);  // along each row (dim 2)
    const double sqF = sqrt(sumF2);
    for (int e = t; e < nE; e += C::TPI) {
      const int s0 = w0, s1 = w1;
      step();
      // sat_window_add order (1-based corners T = s0 + 1, L = s1 + 1)
      double W = 0.0;
      W += satl[(s0 + is0) + Dl * (s1 + is1)];
      W += satl[s0 + Dl * s1];
      W -= satl[s0 + Dl * (s1 + is1)];
      W -= satl[(s0 + is0) + Dl * s1];
      double den = sqrt(W);
      den *= sqF;
      const double r = (double)S[e] / den;  // zncc_ops.jl:39-48
      S[e] = (T)r;
      visit(e, s0, s1, (T)r);
    }
  } else if constexpr (ALG == ALG_ZNCC || ALG == ALG_NSQECC) {
    // W(s) = sum of G'^2 over the search-buffer window [s, s + is) (the SAT
    // window sums of common_ops.jl:203-245), from the image's FP64 integral
    // arrays after clipping the window to the part that holds image data.
    // Each lane takes NB elements per round and issues all their corner loads
    // before using any.  Corner indices are clamped so that an empty window
    // (no image data) reads four equal corners and sums to exactly 0, and
    // elements past the end read element 0's corners; no load is predicated.
    // Corners are addressed as 32-bit byte offsets from the table entry at the
    // clipped search region's first pixel (the host keeps (dims[0] + 2) * 1024
    // < 2^32, and a window spans at most 128 table rows).
    const int64_t D0 = d0 + 1;
    const double sqF = sqrt(sumF2);
    const double mu = (dt == DT_F32) ? (double)meanG32 : meanG64;
    const int64_t corner0 = (g0.sr + D0 * g1.sr) * (int64_t)sizeof(double);
    const gptr<char> sat2 = as_global<char>(J.sat2) + corner0;
    const gptr<char> sat1 = as_global<char>(J.sat1) + corner0;
    const uint32_t rowb = (uint32_t)D0 * (uint32_t)sizeof(double);
    auto at = [](gptr<char> b, uint32_t off) { return *(gptr<double>)(b + off); };
    constexpr int NB = C::TPI >= 384 ? 2 : 4;  // wide IAs: fewer elements per lane, fewer live registers
    constexpr int NQ = (ALG == ALG_ZNCC) ? 2 : 1;
    for (int e0 = t; e0 < nE; e0 += NB * C::TPI) {
      double q[NB][NQ][4];
      double nbox[NB];
      int pw0[NB], pw1[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int e = e0 + j * C::TPI;
        pw0[j] = w0;
        pw1[j] = w1;
        step();
        const int s0 = e < nE ? pw0[j] : 0, s1 = e < nE ? pw1[j] : 0;
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
        q[j][0][0] = at(sat2, c11);
        q[j][0][1] = at(sat2, c01);
        q[j][0][2] = at(sat2, c10);
        q[j][0][3] = at(sat2, c00);
        if constexpr (ALG == ALG_ZNCC) {
          q[j][NQ - 1][0] = at(sat1, c11);
          q[j][NQ - 1][1] = at(sat1, c01);
          q[j][NQ - 1][2] = at(sat1, c10);
          q[j][NQ - 1][3] = at(sat1, c00);
        }
        nbox[j] = (double)((hi0w - lo0w) * (hi1w - lo1w));
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int e = e0 + j * C::TPI;
        if (e >= nE) break;
        double W = ((q[j][0][0] - q[j][0][1]) - q[j][0][2]) + q[j][0][3];
        if constexpr (ALG == ALG_ZNCC) {
          const double S1 = ((q[j][NQ - 1][0] - q[j][NQ - 1][1]) - q[j][NQ - 1][2]) + q[j][NQ - 1][3];
          W = (W - 2.0 * mu * S1) + nbox[j] * mu * mu;
        }
        W = W > 0.0 ? W : 0.0;
        const double c = (double)S[e];
        const double den = sqrt(W) * sqF;
        double r;
        if constexpr (ALG == ALG_ZNCC) {
          r = c / den;  // zncc_ops.jl:39-48
        } else {
          const double num = (sumF2 - 2.0 * c) + W;  // nsqecc_ops.jl:58-60
          r = den / (num + den);                     // nsqecc_ops.jl:72-73
        }
        S[e] = (T)r;
        visit(e, pw0[j], pw1[j], (T)r);
      }
    }
  } else if constexpr (ALG == ALG_MASK_NSQECC) {
    const double sqF = sqrt(sumF2);
    // kMaskRefine: an FP32 transform moves a correlation value by about eps |a|
    // |b| (its two windows' norms), so num by eps (2 |F m| |G| + |m| |G^2|) =
    // eps E12 and den by den eps E2 / (2 C2); the ratio den / (num + den) then
    // moves by eps times fr below, which is kept in C1 for the flag test
    double E12 = 0.0, E2 = 0.0;
    if constexpr (kMaskRefine) {
      if (refine_out) {
        double a = e_fm, b = e_g2, c = e_m, d = e_g4;
        Group<C::TPI>::reduce2(a, b, OpAdd{}, OpAdd{}, red, t);
        Group<C::TPI>::reduce2(c, d, OpAdd{}, OpAdd{}, red, t);
        E2 = sqrt(c * d);
        E12 = 2.0 * sqrt(a * b) + E2;
      }
    }
    for (int i = t; i < nE; i += C::TPI) {
      const double c1 = (double)C1[i];
      const double c2 = (double)S[i];
      const double num = (-2.0 * c1 + sumF2) + c2;  // masked_nsqecc_ops.jl:130-132
      const double den = sqrt(fabs(c2)) * sqF;      // masked_nsqecc_ops.jl:135-136
      S[i] = (T)(den / (num + den));
      if constexpr (kMaskRefine) {
        if (refine_out) {
          const double q = num + den;
          C1[i] = (T)((den * E12 + num * den * E2 / (2.0 * fabs(c2))) / (q * q));
        }
      }
      visit(i, w0, w1, (T)(den / (num + den)));
      step();
    }
  } else {
    for (int i = t; i < nE; i += C::TPI) {
      visit(i, w0, w1, S[i]);
      step();
    }
  }
  // combined reduction: (peak value, index), SN maximum, SN sum; its barrier
  // also makes the normalised window visible to the fit below
  {
    // (idle partners carry the identities: value 0 with no index, -inf, 0)
    Group<C::TPI>::steps([&](auto mv) {
      const Best<T> o{mv(T(0), bst.v), mv((int32_t)INT32_MAX, bst.i)};
      bst = best_of(bst, o);
      mx = OpMaxNaN{}(mx, mv(OpMaxNaN::id(), mx));
      smv += mv(0.0, smv);
    });
    if constexpr (C::TPI == 64) {
      bst.v = bcast63(bst.v);
      bst.i = bcast63(bst.i);
      mx = bcast63(mx);
      smv = bcast63(smv);
    }
    __syncthreads();
    if constexpr (C::TPI > 64) {
      if ((t & 63) == 63) {
        const int w = t >> 6;
        red[w] = (double)bst.v;
        red[16 + w] = (double)bst.i;
        red[32 + w] = mx;
        red[48 + w] = smv;
      }
      __syncthreads();
      Best<T> r{(T)red[0], (int32_t)red[16]};
      double rm = red[32], rs = red[48];
#pragma unroll 4
      for (int w = 1; w < C::NW; ++w) {
        r = best_of(r, Best<T>{(T)red[w], (int32_t)red[16 + w]});
        rm = OpMaxNaN{}(rm, red[32 + w]);
        rs += red[48 + w];
      }
      bst = r;
      mx = rm;
      smv = rs;
    }
  }
  double sn = 0.0;
  const int u0 = shi0 - lo0 + 1, u1 = shi1 - lo1 + 1;
  const int nsn = (u0 > 0 && u1 > 0) ? u0 * u1 : 0;
  if (want_sn) sn = nsn > 0 ? mx / ((smv - mx) / (double)(nsn - 1)) : (double)NAN;
  // kMaskRefine: the runner-up of the peak window (the largest value at another
  // position; +inf when a value there is not finite) and the sum of fr over the
  // SN window
  double t2 = -INFINITY, sfr = 0.0;
  if constexpr (kMaskRefine) {
    if (refine_out) {
      int a0 = t % ns0, a1 = t / ns0;
      for (int e = t; e < nE; e += C::TPI) {
        if (a0 >= lo0 && a0 <= hi0 && a1 >= lo1 && a1 <= hi1) {
          const double v = (double)S[e];
          if (!(__builtin_fabs(v) <= 1.7976931348623157e308)) t2 = INFINITY;
          else if (e != bst.i && v > t2) t2 = v;
        }
        if (a0 >= lo0 && a0 <= shi0 && a1 >= lo1 && a1 <= shi1) sfr += (double)C1[e];
        a0 += st0;
        a1 += st1;
        if (a0 >= ns0) {
          a0 -= ns0;
          ++a1;
        }
      }
      Group<C::TPI>::reduce2(t2, sfr, OpMax{}, OpAdd{}, red, t);
    }
  }

  // ---- Gaussian subpixel + write (piv_ops.jl:249-326, 236-243).  The 3x3
  // neighbourhood is spread over the IA's first 9 lanes (lane 3*(b1+1)+(b0+1)
  // holds offset (b0, b1)); its minimum is a shuffle reduction and each lane
  // takes the log of its own value, so the five logs run side by side; lane 0
  // combines them.  Same operations as a serial evaluation, bit for bit.
  const bool has_peak = !skip && bst.i != INT32_MAX;
  const int pk0 = has_peak ? bst.i % ns0 : 0, pk1 = has_peak ? bst.i / ns0 : 0;  // 0-based
  const bool interior = has_peak && (pk0 > lo0) && (pk0 < hi0) && (pk1 > lo1) && (pk1 < hi1);
  double lfit[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // lane 0: log(1 + v - mn) at centre, up, down, left, right
  if (interior && t < 64) {  // uniform over the IA's first wave; other waves skip
    const int tt = t < 9 ? t : 4;  // lanes past 8 repeat the centre (neutral for the minimum)
    const double v = (double)S[(pk0 + tt % 3 - 1) + ns0 * (pk1 + tt / 3 - 1)];
    double mn = v;
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) {
      const double o = __shfl_xor(mn, m);
      mn = (mn != mn || o != o) ? (double)NAN : (o < mn ? o : mn);
    }
    const double lv = log(1.0 + v - mn);
    const int lane0 = (int)(threadIdx.x & 63) - (t & 63);  // lane of t = 0 in this wave
    lfit[0] = __shfl(lv, lane0 + 4);
    lfit[1] = __shfl(lv, lane0 + 3);
    lfit[2] = __shfl(lv, lane0 + 5);
    lfit[3] = __shfl(lv, lane0 + 1);
    lfit[4] = __shfl(lv, lane0 + 7);
  }
  if (valid && t == 0) {
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
        const double q0 = (up - dn) / (2 * up - 4 * lmx + 2 * dn);
        const double q1 = (lf - rt) / (2 * lf - 4 * lmx + 2 * rt);
        r0 = ((double)(pk0 + 1) + q0) - (double)(sm0 + 1);
        r1 = ((double)(pk1 + 1) + q1) - (double)(sm1 + 1);
      }
    }
    if (SH && !skip) {  // the displacement is measured from the shifted search region
      r0 += (double)sh0;
      r1 += (double)sh1;
    }
    if constexpr (kMaskRefine) {
      // (an IA whose IA mask is all 0 has den = 0 and only zeros in both forms)
      if (refine_out && !skip && sumF2 > 0.0) {
        bool f = false;
        if (has_peak) {
          const double fpk = (double)C1[bst.i];
          f |= !((double)bst.v - t2 >= kMaskRefinePeak * fpk);
          if (interior) {
            const double lmx = lfit[0], up = lfit[1], dn = lfit[2], lf = lfit[3], rt = lfit[4];
            const double D0 = 2 * up - 4 * lmx + 2 * dn, D1 = 2 * lf - 4 * lmx + 2 * rt;
            const double e = kMaskRefineEps * 10.0 * fpk;  // log error, ten times the value error bound
            f |= !(1e3 * e * (2.0 + 8.0 * __builtin_fabs((up - dn) / D0)) <= __builtin_fabs(D0)) ||
                 !(1e3 * e * (2.0 + 8.0 * __builtin_fabs((lf - rt) / D1)) <= __builtin_fabs(D1));
          }
        }
        if (want_sn && nsn > 1) {
          const double avg = (smv - mx) / (double)(nsn - 1);
          f |= !(kMaskRefineEps * (sfr / (double)nsn) <= kMaskRefineSn * __builtin_fabs(avg));
        }
        if (f) {
          const int32_t pos = atomicAdd(refine_out, 1);
          refine_out[kRefineHead + pos] = ia_g;
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
  if constexpr (GZ) {
    __syncthreads();  // every wave is done with the buffer
    if (tid == 0) {
      __threadfence();
      atomicExch(reinterpret_cast<int32_t*>(gz) + gz_slot, 0);
    }
  }
  if constexpr (!LIST) break;
  lpos += (int32_t)gridDim.x;
  __syncthreads();  // the next IA rewrites the LDS buffers
  }
}


// ---------------------------------------------------------------------------
// Host-side launch helpers (explicitly instantiated per (T, L) in mqpiv_k2d_*.hip).

template <int L, typename T, int LPL>
size_t lds_bytes_2d(int alg, int ns0, int ns1, int ss0) {
  using C = Cfg2<L, T, LPL>;
  size_t z = gz_len<L, T>() ? 0 : region_bytes_2d(L, sizeof(T), ns0, ns1, ss0);
  // FP64 ZNCC with a global transform buffer keeps its window-energy table in LDS
  const size_t zw = (gz_len<L, T>() && alg == ALG_ZNCC) ? ((size_t)(L + 1) * (L + 1) + 1) * sizeof(double) : 0;
  size_t c1 = (alg == ALG_MASK_NSQECC) ? sim_bytes(ns0, ns1, sizeof(T)) : 0;
  return (z + c1) * C::IPB + (size_t)C::IPB * kRedBytes + zw;
}

template <int L, int ALG, typename T, int LPL, bool SH>
hipError_t launch_2d_k(const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs, int64_t n_ia_total,
                       int ns0, int ns1, size_t lds, char* gz, hipStream_t s, int32_t* refine_out) {
  using C = Cfg2<L, T, LPL>;
  const void* fn = (const void*)piv2d_kernel<L, ALG, T, LPL, SH, false>;
  allow_full_lds(fn);
  const int64_t nb = (n_ia_total + C::IPB - 1) / C::IPB;
  int slots = 0;
  if constexpr (gz_len<L, T>()) {
    if (!gz) return hipErrorInvalidValue;  // the caller provides gz_scratch_2d bytes
    slots = gz_slots_2d(n_ia_total);
  }
  hipLaunchKernelGGL((piv2d_kernel<L, ALG, T, LPL, SH, false>), dim3((unsigned)nb), dim3(C::NT), lds, s, j0, d_jobs,
                     d_prefix, n_jobs, n_ia_total, ns0, ns1, gz, slots, (const int32_t*)nullptr, refine_out);
  return hipGetLastError();
}

// ZNCC or masked NSQECC over the IAs of a refine list (refine_list layout)
// that an earlier kernel on the stream wrote: a persistent grid of as many
// blocks as the transform buffers (LDS, or for gz_len lengths the global slots
// in gz) let the CUs hold, each stepping through the list.
template <int L, int ALG, typename T, int LPL>
hipError_t launch_2d_refine_a(const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs,
                              int64_t n_ia_total, int ns0, int ns1, int ss0, const int32_t* list, char* gz,
                              hipStream_t s) {
  using C = Cfg2<L, T, LPL>;
  const void* fn = (const void*)piv2d_kernel<L, ALG, T, LPL, false, true>;
  allow_full_lds(fn);
  const size_t lds = lds_bytes_2d<L, T, LPL>(ALG, ns0, ns1, ss0);
  if (lds > 160 * 1024 || !list) return hipErrorInvalidValue;
  int64_t per_cu;
  int slots = 0;
  if constexpr (gz_len<L, T>()) {
    if (!gz) return hipErrorInvalidValue;  // the caller provides gz_scratch_2d bytes
    per_cu = kGzBlocksPerCu;
    slots = gz_slots_2d(n_ia_total);
  } else {
    const int64_t by_lds = (int64_t)((160 * 1024) / (lds > 0 ? lds : 1));
    const int64_t by_waves = (int64_t)min_waves_2d<L, ALG, T, LPL>() * 4 * 64 / C::NT;
    per_cu = by_lds < by_waves ? by_lds : by_waves;
  }
  if (per_cu < 1) per_cu = 1;
  int64_t nb = (n_ia_total + C::IPB - 1) / C::IPB;
  if (nb > per_cu * device_cus()) nb = per_cu * device_cus();
  if (gz_len<L, T>() && nb > slots) nb = slots;  // every resident block finds a free slot
  if (nb < 1) return hipSuccess;
  hipLaunchKernelGGL((piv2d_kernel<L, ALG, T, LPL, false, true>), dim3((unsigned)nb), dim3(C::NT), lds, s, j0,
                     d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, gz, slots, list, (int32_t*)nullptr);
  return hipGetLastError();
}

template <int L, typename T, int LPL>
hipError_t launch_2d_refine(int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs,
                            int64_t n_ia_total, int ns0, int ns1, int ss0, const int32_t* list, char* gz,
                            hipStream_t s) {
  if (alg == ALG_ZNCC)
    return launch_2d_refine_a<L, ALG_ZNCC, T, LPL>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, list, gz,
                                                  s);
  if (alg == ALG_MASK_NSQECC)
    return launch_2d_refine_a<L, ALG_MASK_NSQECC, T, LPL>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0,
                                                         list, gz, s);
  return hipErrorInvalidValue;
}

// A single-job launch whose job carries search-region shifts runs the SH
// instantiation (batches never carry shifts).
template <int L, int ALG, typename T, int LPL>
hipError_t launch_2d_t(const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs,
                       int64_t n_ia_total, int ns0, int ns1, int ss0, char* gz, hipStream_t s, int32_t* refine_out) {
  size_t lds = lds_bytes_2d<L, T, LPL>(ALG, ns0, ns1, ss0);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  if (refine_out && (j0.shift || ALG != ALG_MASK_NSQECC || !std::is_same<T, float>::value))
    return hipErrorInvalidValue;  // FP32 masked NSQECC flags IAs, unshifted
  if (j0.shift && !d_jobs)
    return launch_2d_k<L, ALG, T, LPL, true>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, lds, gz, s, nullptr);
  return launch_2d_k<L, ALG, T, LPL, false>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, lds, gz, s,
                                            refine_out);
}

template <int L, typename T, int LPL>
hipError_t launch_2d_l(int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix, int n_jobs,
                       int64_t n_ia_total, int ns0, int ns1, int ss0, char* gz, hipStream_t s, int32_t* refine_out) {
  if (refine_out && alg != ALG_MASK_NSQECC) return hipErrorInvalidValue;
  switch (alg) {
    case ALG_FFTCC:
      return launch_2d_t<L, ALG_FFTCC, T, LPL>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, gz, s,
                                               nullptr);
    case ALG_ZNCC:
      return launch_2d_t<L, ALG_ZNCC, T, LPL>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, gz, s, nullptr);
    case ALG_NSQECC:
      return launch_2d_t<L, ALG_NSQECC, T, LPL>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1, ss0, gz, s,
                                                nullptr);
    case ALG_MASK_NSQECC:
      return launch_2d_t<L, ALG_MASK_NSQECC, T, LPL>(j0, d_jobs, d_prefix, n_jobs, n_ia_total, ns0, ns1,
                                                    ss0, gz, s, refine_out);
  }
  return hipErrorInvalidValue;
}

}  // namespace mqpiv
