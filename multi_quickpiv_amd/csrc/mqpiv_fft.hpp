// mqpiv_fft.hpp -- register-resident complex FFTs for gfx950.
//
// A line of L complex samples is transformed in VGPRs by one lane or split over
// an aligned group of 2 or 4 lanes: a self-sorting (Stockham) sequence of
// radix-4/3/2 passes, fully unrolled, with every twiddle factor a compile-time
// constant.  A lane group runs decimation in time in the forward direction and
// decimation in frequency in the inverse direction, with radix-2 stages across
// the group exchanged through DPP quad permutations.  A 2-D transform is two
// such sweeps with an LDS transpose in between; see mqpiv_kernels2d.hpp.
//
// The transforms replace FFTW's in-place r2c/c2r pair that the reference drives
// through its C API (src/utils/fft_ops.jl:11-83).  The lengths used here are the
// smallest supported length >= the search size, which yields the same
// fully-overlapping correlation values as the reference's own transform size (no
// circular wrap reaches the window [0, 2*SM]).
#pragma once
#include <hip/hip_runtime.h>

namespace mqpiv {

// ---------------------------------------------------------------------------
// Compile-time cos/sin of 2*pi*k/L: exact integer reduction of k to [-L/2, L/2],
// then a Taylor series on |angle| <= pi (error < 1e-16).
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double taylor_sin(double x) {
  double t = x, s = x;
  for (int n = 1; n < 16; ++n) {
    t *= -x * x / double((2 * n) * (2 * n + 1));
    s += t;
  }
  return s;
}
constexpr double taylor_cos(double x) {
  double t = 1.0, s = 1.0;
  for (int n = 1; n < 16; ++n) {
    t *= -x * x / double((2 * n - 1) * (2 * n));
    s += t;
  }
  return s;
}
constexpr long red_k(long k, long L) {
  k %= L;
  if (k < 0) k += L;
  if (2 * k > L) k -= L;
  return k;
}
constexpr double twc(long k, long L) { return taylor_cos(2.0 * kPi * double(red_k(k, L)) / double(L)); }
constexpr double tws(long k, long L) { return taylor_sin(2.0 * kPi * double(red_k(k, L)) / double(L)); }

template <int L, typename T>
struct Twiddles {
  T c[L];
  T s[L];
  constexpr Twiddles() : c(), s() {
    for (int k = 0; k < L; ++k) {
      c[k] = T(twc(k, L));
      s[k] = T(tws(k, L));
    }
  }
};

// ---------------------------------------------------------------------------
template <typename T>
struct cpx {
  T x, y;
};
using cf = cpx<float>;
using cd = cpx<double>;

template <typename T>
__host__ __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
  return {a.x - b.x, a.y - b.y};
}

// u * exp(SIGN * 2*pi*i * idx / L); SIGN = -1 forward, +1 inverse.
template <int L, int SIGN, typename T>
__host__ __device__ __forceinline__ cpx<T> twiddle_mul(cpx<T> u, int idx) {
  constexpr Twiddles<L, T> W{};
  idx %= L;
  if (idx == 0) return u;
  if (4 * idx == L) return SIGN > 0 ? cpx<T>{-u.y, u.x} : cpx<T>{u.y, -u.x};
  if (2 * idx == L) return cpx<T>{-u.x, -u.y};
  if (4 * idx == 3 * L) return SIGN > 0 ? cpx<T>{u.y, -u.x} : cpx<T>{-u.y, u.x};
  const T c = W.c[idx];
  const T s = SIGN > 0 ? W.s[idx] : -W.s[idx];
  return cpx<T>{u.x * c - u.y * s, u.x * s + u.y * c};
}

template <int R, int SIGN, typename T>
struct Butterfly;

template <int SIGN, typename T>
struct Butterfly<2, SIGN, T> {
  __host__ __device__ __forceinline__ static void run(cpx<T> (&v)[2]) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int SIGN, typename T>
struct Butterfly<4, SIGN, T> {
  __host__ __device__ __forceinline__ static void run(cpx<T> (&v)[4]) {
    cpx<T> a0 = cadd(v[0], v[2]);
    cpx<T> a1 = csub(v[0], v[2]);
    cpx<T> a2 = cadd(v[1], v[3]);
    cpx<T> a3 = csub(v[1], v[3]);
    cpx<T> b3 = SIGN > 0 ? cpx<T>{-a3.y, a3.x} : cpx<T>{a3.y, -a3.x};  // a3 * (SIGN i)
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, b3);
    v[3] = csub(a1, b3);
  }
};

template <int SIGN, typename T>
struct Butterfly<3, SIGN, T> {
  __host__ __device__ __forceinline__ static void run(cpx<T> (&v)[3]) {
    // y_k = sum_q v_q exp(SIGN*2*pi*i*k*q/3)
    const T c1 = T(-0.5);
    const T s1 = T(SIGN) * T(0.86602540378443864676372317075294);
    cpx<T> t = cadd(v[1], v[2]);
    cpx<T> d = csub(v[1], v[2]);
    cpx<T> y0 = cadd(v[0], t);
    cpx<T> m = {v[0].x + c1 * t.x, v[0].y + c1 * t.y};
    cpx<T> r = {-s1 * d.y, s1 * d.x};  // i * s1 * d
    v[0] = y0;
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

// One Stockham pass of radix R over L samples; NS = product of the radices
// already applied.  Reads a[], writes b[] in self-sorted order.
template <int L, int R, int NS, int SIGN, typename T>
__host__ __device__ __forceinline__ void stockham_pass(const cpx<T> (&a)[L], cpx<T> (&b)[L]) {
  constexpr int M = L / R;
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const int jn = j % NS;
    cpx<T> v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = twiddle_mul<L, SIGN>(a[j + q * M], jn * q * (L / (NS * R)));
    Butterfly<R, SIGN, T>::run(v);
    const int idxD = (j / NS) * NS * R + jn;
#pragma unroll
    for (int q = 0; q < R; ++q) b[idxD + q * NS] = v[q];
  }
}

// Radix plans: the product of the listed radices is L.
template <int L>
struct Plan;
template <> struct Plan<4> { static constexpr int n = 1; static constexpr int r[4] = {4, 1, 1, 1}; };
template <> struct Plan<6> { static constexpr int n = 2; static constexpr int r[4] = {3, 2, 1, 1}; };
template <> struct Plan<8> { static constexpr int n = 2; static constexpr int r[4] = {4, 2, 1, 1}; };
template <> struct Plan<12> { static constexpr int n = 2; static constexpr int r[4] = {4, 3, 1, 1}; };
template <> struct Plan<16> { static constexpr int n = 2; static constexpr int r[4] = {4, 4, 1, 1}; };
template <> struct Plan<24> { static constexpr int n = 3; static constexpr int r[4] = {4, 3, 2, 1}; };
template <> struct Plan<32> { static constexpr int n = 3; static constexpr int r[4] = {4, 4, 2, 1}; };
template <> struct Plan<48> { static constexpr int n = 3; static constexpr int r[4] = {4, 4, 3, 1}; };
template <> struct Plan<64> { static constexpr int n = 3; static constexpr int r[4] = {4, 4, 4, 1}; };
template <> struct Plan<96> { static constexpr int n = 4; static constexpr int r[4] = {4, 4, 2, 3}; };
template <> struct Plan<128> { static constexpr int n = 4; static constexpr int r[4] = {4, 4, 4, 2}; };

// In-place transform of one line held entirely by the calling lane.
template <int L, int SIGN, typename T>
__host__ __device__ __forceinline__ void fft_line(cpx<T> (&x)[L]) {
  cpx<T> y[L];
  constexpr int n = Plan<L>::n;
  constexpr int r0 = Plan<L>::r[0];
  constexpr int r1 = Plan<L>::r[1];
  constexpr int r2 = Plan<L>::r[2];
  constexpr int r3 = Plan<L>::r[3];
  stockham_pass<L, r0, 1, SIGN>(x, y);
  if constexpr (n == 1) {
#pragma unroll
    for (int i = 0; i < L; ++i) x[i] = y[i];
  } else if constexpr (n == 2) {
    stockham_pass<L, r1, r0, SIGN>(y, x);
  } else if constexpr (n == 3) {
    stockham_pass<L, r1, r0, SIGN>(y, x);
    stockham_pass<L, r2, r0 * r1, SIGN>(x, y);
#pragma unroll
    for (int i = 0; i < L; ++i) x[i] = y[i];
  } else {
    stockham_pass<L, r1, r0, SIGN>(y, x);
    stockham_pass<L, r2, r0 * r1, SIGN>(x, y);
    stockham_pass<L, r3, r0 * r1 * r2, SIGN>(y, x);
  }
}

// ---------------------------------------------------------------------------
// Lane-group transforms: a line of length L split over P = 2, 4 or 8 adjacent
// lanes (an aligned group), each holding M = L/P samples.
//
// Forward (decimation in time): on entry lane q holds x[P*k + q], k < M; each
// lane transforms its decimated subsequence, then log2(P) radix-2 stages across
// the group (partners q ^ P/2, ..., q ^ 2, q ^ 1) combine them; on exit lane q
// holds X[chunk(q)*M + k] with chunk(q) = the log2(P)-bit reversal of q.
// Inverse direction (decimation in frequency): on entry lane q holds
// x[chunk(q)*M + k]; the radix-2 stages run first in the reverse order
// (q ^ 1, ..., q ^ P/2), then the local transform; on exit lane q holds
// X[P*m + q].
//
// Stage with partner mask 2^j (forward order j = log2(P)-1 .. 0) works on
// sub-transforms of length L / 2^j; the twiddle offset of a lane is
// chunk_{P/2^(j+1)}(q >> (j+1)) * M (the chunk its sub-group produced).
//
// The partner exchange is a callable so the same code runs on the CPU in unit
// tests.  On the GPU, q ^ 1 and q ^ 2 are DPP quad permutations (one VALU move
// per dword) and q ^ 4, q ^ 8 are ds_swizzle xors.

template <int P>
__host__ __device__ constexpr int chunk_of(int q) {
  int r = 0;
  for (int b = 1, c = P >> 1; b < P; b <<= 1, c >>= 1)
    if (q & b) r |= c;
  return r;
}

#if defined(__HIP_DEVICE_COMPILE__)
template <int MASK>
__device__ __forceinline__ int xor_lane_i32(int v) {
  if constexpr (MASK == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
  else if constexpr (MASK == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
  else return __builtin_amdgcn_ds_swizzle(v, (MASK << 10) | 0x1F);  // bit mode: xor MASK
}
template <int MASK>
__device__ __forceinline__ float xor_lane(float v) {
  return __builtin_bit_cast(float, xor_lane_i32<MASK>(__builtin_bit_cast(int, v)));
}
template <int MASK>
__device__ __forceinline__ double xor_lane(double v) {
  long long u = __builtin_bit_cast(long long, v);
  int lo = xor_lane_i32<MASK>((int)(u & 0xffffffffll));
  int hi = xor_lane_i32<MASK>((int)(u >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
#endif

// Device exchange with lane q ^ MASK of the same group.
struct DppXchg {
  template <int MASK, typename T>
  __device__ __forceinline__ cpx<T> get(cpx<T> v, int) const {
#if defined(__HIP_DEVICE_COMPILE__)
    return cpx<T>{xor_lane<MASK>(v.x), xor_lane<MASK>(v.y)};
#else
    return v;
#endif
  }
};

// Fused multiply-add in the element type (host and device).
template <typename T>
__host__ __device__ __forceinline__ T fma_t(T a, T b, T c) {
  if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

// u * -exp(SIGN * 2*pi*i * K / L) for a compile-time K: the negated twiddle
// that the high member of a cross-lane stage applies.
template <int L, int SIGN, int K, typename T>
__host__ __device__ __forceinline__ cpx<T> twiddle_mul_neg(cpx<T> u) {
  constexpr int k = K % L;
  if constexpr (k == 0) return cpx<T>{-u.x, -u.y};
  else if constexpr (4 * k == L) return SIGN > 0 ? cpx<T>{u.y, -u.x} : cpx<T>{-u.y, u.x};
  else if constexpr (2 * k == L) return u;
  else if constexpr (4 * k == 3 * L) return SIGN > 0 ? cpx<T>{-u.y, u.x} : cpx<T>{u.y, -u.x};
  else {
    constexpr Twiddles<L, T> W{};
    constexpr T c = -W.c[k];
    constexpr T s = SIGN > 0 ? -W.s[k] : W.s[k];
    return cpx<T>{u.x * c - u.y * s, u.x * s + u.y * c};
  }
}

// The lane-dependent factor of a cross-lane stage.  A stage's twiddle is
// W_LEN^(k + c*M) with k < M compile-time and c in [0, NC) the lane's chunk;
// since LEN = 2*NC*M it factors as W_LEN^k * exp(SIGN*i*pi*c/NC): NC = 1 needs
// nothing, NC = 2 a rotation by SIGN*i (selects), NC = 4 or 8 one runtime
// complex product with a factor picked once per stage (NC = 8: the NC = 4
// factor of c/2 times exp(SIGN*i*pi/8) for odd c).  No runtime-indexed tables.
template <int NC, int SIGN, typename T>
__host__ __device__ __forceinline__ cpx<T> lane_factor(int c) {
  if constexpr (NC == 4) {
    constexpr T h = T(0.70710678118654752440084436210485);
    const bool b0 = (c & 1) != 0, b1 = (c & 2) != 0;
    const T re = b1 ? (b0 ? -h : T(0)) : (b0 ? h : T(1));
    const T im = b1 ? (b0 ? h : T(1)) : (b0 ? h : T(0));
    return cpx<T>{re, SIGN > 0 ? im : -im};
  } else if constexpr (NC == 8) {
    const cpx<T> f = lane_factor<4, SIGN, T>(c >> 1);
    constexpr T co = T(0.92387953251128675612818318939678829);  // cos(pi/8)
    constexpr T si = T(SIGN > 0 ? 0.38268343236508977172845998403039887 : -0.38268343236508977172845998403039887);
    if ((c & 1) == 0) return f;
    return cpx<T>{f.x * co - f.y * si, f.x * si + f.y * co};
  } else {
    return cpx<T>{T(1), T(0)};
  }
}
template <int NC, int SIGN, typename T>
__host__ __device__ __forceinline__ cpx<T> apply_lane_factor(cpx<T> u, int c, cpx<T> f) {
  if constexpr (NC == 1) {
    return u;
  } else if constexpr (NC == 2) {
    // u * (SIGN*i)^c
    const cpx<T> r = SIGN > 0 ? cpx<T>{-u.y, u.x} : cpx<T>{u.y, -u.x};
    return c ? r : u;
  } else {
    return cpx<T>{u.x * f.x - u.y * f.y, u.x * f.y + u.y * f.x};
  }
}

template <int LEN, int SIGN, int NC, int M, int K, typename T>
__host__ __device__ __forceinline__ void hi_twiddle(cpx<T> (&v)[M], int c, cpx<T> f) {
  if constexpr (K < M) {
    v[K] = apply_lane_factor<NC, SIGN>(twiddle_mul_neg<LEN, SIGN, K>(v[K]), c, f);
    hi_twiddle<LEN, SIGN, NC, M, K + 1>(v, c, f);
  }
}

// One radix-2 stage across lanes q and q ^ MASK (DIT form).  The high member
// (q & MASK) holds the odd half and replaces it by -O*W (negated twiddle); the
// exchange then gives both members the same form, out = self + s * partner
// with s = -1 (low: E + O*W) or +1 (high: E - O*W), a single fused multiply-add
// per component with the partner value straight from the cross-lane move.
template <int LEN, int SIGN, int MASK, int NC, int M, typename T, typename X>
__host__ __device__ __forceinline__ void group_dit_stage(cpx<T> (&v)[M], int q, int c, const X& xchg) {
  const bool hi = (q & MASK) != 0;
  if (hi) hi_twiddle<LEN, SIGN, NC, M, 0>(v, c, lane_factor<NC, SIGN, T>(c));
  const T s = hi ? T(1) : T(-1);
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const cpx<T> u = xchg.template get<MASK>(v[k], k);
    v[k] = cpx<T>{fma_t(u.x, s, v[k].x), fma_t(u.y, s, v[k].y)};
  }
}

// One radix-2 stage across lanes q and q ^ MASK (DIF form): out = self +
// s * partner with s = +1 (low: a + b) or -1 (high: b - a), then the high
// member multiplies by the negated twiddle, giving (a - b) * W.
template <int LEN, int SIGN, int MASK, int NC, int M, typename T, typename X>
__host__ __device__ __forceinline__ void group_dif_stage(cpx<T> (&v)[M], int q, int c, const X& xchg) {
  const bool hi = (q & MASK) != 0;
  const T s = hi ? T(-1) : T(1);
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const cpx<T> u = xchg.template get<MASK>(v[k], k);
    v[k] = cpx<T>{fma_t(u.x, s, v[k].x), fma_t(u.y, s, v[k].y)};
  }
  if (hi) hi_twiddle<LEN, SIGN, NC, M, 0>(v, c, lane_factor<NC, SIGN, T>(c));
}

// Stage j of a P-lane group (partner mask 2^j).
template <int L, int P, int J, int SIGN, bool DIT, int M, typename T, typename X>
__host__ __device__ __forceinline__ void group_stage(cpx<T> (&v)[M], int q, const X& xchg) {
  constexpr int MASK = 1 << J;
  constexpr int LEN = L >> J;
  constexpr int NC = P >> (J + 1);
  const int c = chunk_of<NC>(q >> (J + 1));
  if constexpr (DIT) group_dit_stage<LEN, SIGN, MASK, NC, M>(v, q, c, xchg);
  else group_dif_stage<LEN, SIGN, MASK, NC, M>(v, q, c, xchg);
}

// Forward transform over a group of P lanes (see above).
template <int L, int P, int SIGN, typename T, typename X>
__host__ __device__ __forceinline__ void fft_group_dit(cpx<T> (&v)[L / P], int q, const X& xchg) {
  constexpr int M = L / P;
  static_assert(P == 2 || P == 4 || P == 8 || P == 16, "groups of 2, 4, 8 or 16 lanes");
  fft_line<M, SIGN, T>(v);
  if constexpr (P >= 16) group_stage<L, P, 3, SIGN, true, M>(v, q, xchg);
  if constexpr (P >= 8) group_stage<L, P, 2, SIGN, true, M>(v, q, xchg);
  if constexpr (P >= 4) group_stage<L, P, 1, SIGN, true, M>(v, q, xchg);
  group_stage<L, P, 0, SIGN, true, M>(v, q, xchg);
}

// Inverse-direction transform over a group of P lanes (see above).
template <int L, int P, int SIGN, typename T, typename X>
__host__ __device__ __forceinline__ void fft_group_dif(cpx<T> (&v)[L / P], int q, const X& xchg) {
  constexpr int M = L / P;
  static_assert(P == 2 || P == 4 || P == 8 || P == 16, "groups of 2, 4, 8 or 16 lanes");
  group_stage<L, P, 0, SIGN, false, M>(v, q, xchg);
  if constexpr (P >= 4) group_stage<L, P, 1, SIGN, false, M>(v, q, xchg);
  if constexpr (P >= 8) group_stage<L, P, 2, SIGN, false, M>(v, q, xchg);
  if constexpr (P >= 16) group_stage<L, P, 3, SIGN, false, M>(v, q, xchg);
  fft_line<M, SIGN, T>(v);
}

}  // namespace mqpiv
