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
  if constexpr (n == 2) {
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
// Lane-group transforms: a line of length L split over P = 2 or 4 adjacent
// lanes (an aligned pair / quad), each holding M = L/P samples.
//
// Forward (decimation in time): on entry lane q holds x[P*k + q], k < M; each
// lane transforms its decimated subsequence, then log2(P) radix-2 stages across
// the group (partner = q ^ 2 first, then q ^ 1 for P = 4) combine them; on exit
// lane q holds X[chunk(q)*M + k] with chunk(q) = bit-reverse of q.
// Inverse direction (decimation in frequency): on entry lane q holds
// x[chunk(q)*M + k]; radix-2 stages across the group (q ^ 1, then q ^ 2) come
// first, then the local transform; on exit lane q holds X[P*m + q].
//
// The partner exchange is a callable so the same code runs on the CPU in unit
// tests.  On the GPU it is a DPP quad permutation (one VALU move per dword).

template <int P>
__host__ __device__ constexpr int chunk_of(int q) {
  return P == 2 ? q : ((q & 1) << 1) | (q >> 1);
}

#if defined(__HIP_DEVICE_COMPILE__)
template <int MASK>
__device__ __forceinline__ float dpp_xor(float v) {
  // quad_perm selectors: lane ^ 1 -> [1,0,3,2] = 0xB1, lane ^ 2 -> [2,3,0,1] = 0x4E
  constexpr int ctrl = MASK == 1 ? 0xB1 : 0x4E;
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), ctrl, 0xF, 0xF, false));
}
template <int MASK>
__device__ __forceinline__ double dpp_xor(double v) {
  constexpr int ctrl = MASK == 1 ? 0xB1 : 0x4E;
  long long u = __builtin_bit_cast(long long, v);
  int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffll), ctrl, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), ctrl, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
#endif

// Device exchange with the lane q ^ MASK of the same quad.
struct DppXchg {
  template <int MASK, typename T>
  __device__ __forceinline__ cpx<T> get(cpx<T> v, int) const {
#if defined(__HIP_DEVICE_COMPILE__)
    return cpx<T>{dpp_xor<MASK>(v.x), dpp_xor<MASK>(v.y)};
#else
    return v;
#endif
  }
};

// One radix-2 stage across lanes q and q ^ MASK (DIT form): the member with
// (q & MASK) != 0 first multiplies by W_len^(k + off), then both exchange;
// the low member keeps E + W*O, the high member E - W*O.
template <int LEN, int SIGN, int MASK, int M, typename T, typename X>
__host__ __device__ __forceinline__ void group_dit_stage(cpx<T> (&v)[M], int q, int off, const X& xchg) {
  const bool hi = (q & MASK) != 0;
  if (hi) {
#pragma unroll
    for (int k = 0; k < M; ++k) v[k] = twiddle_mul<LEN, SIGN>(v[k], k + off);
  }
#pragma unroll
  for (int k = 0; k < M; ++k) {
    cpx<T> u = xchg.template get<MASK>(v[k], k);
    v[k] = hi ? csub(u, v[k]) : cadd(v[k], u);
  }
}

// One radix-2 stage across lanes q and q ^ MASK (DIF form): exchange, the low
// member keeps a + b, the high member (a - b) * W_len^(k + off).
template <int LEN, int SIGN, int MASK, int M, typename T, typename X>
__host__ __device__ __forceinline__ void group_dif_stage(cpx<T> (&v)[M], int q, int off, const X& xchg) {
  const bool hi = (q & MASK) != 0;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    cpx<T> u = xchg.template get<MASK>(v[k], k);
    v[k] = hi ? csub(u, v[k]) : cadd(v[k], u);
  }
  if (hi) {
#pragma unroll
    for (int k = 0; k < M; ++k) v[k] = twiddle_mul<LEN, SIGN>(v[k], k + off);
  }
}

// Forward transform over a group of P lanes (see above).
template <int L, int P, int SIGN, typename T, typename X>
__host__ __device__ __forceinline__ void fft_group_dit(cpx<T> (&v)[L / P], int q, const X& xchg) {
  constexpr int M = L / P;
  fft_line<M, SIGN, T>(v);
  if constexpr (P == 2) {
    group_dit_stage<L, SIGN, 1, M>(v, q, 0, xchg);
  } else {
    static_assert(P == 4, "groups of 2 or 4 lanes");
    // sub-transforms of length L/2 over lanes {q, q^2}
    group_dit_stage<L / 2, SIGN, 2, M>(v, q, 0, xchg);
    // lanes q (even) and q^1 now hold E/O at chunk c = q >> 1 of length-L/2 halves
    group_dit_stage<L, SIGN, 1, M>(v, q, (q >> 1) * M, xchg);
  }
}

// Inverse-direction transform over a group of P lanes (see above).
template <int L, int P, int SIGN, typename T, typename X>
__host__ __device__ __forceinline__ void fft_group_dif(cpx<T> (&v)[L / P], int q, const X& xchg) {
  constexpr int M = L / P;
  if constexpr (P == 2) {
    group_dif_stage<L, SIGN, 1, M>(v, q, 0, xchg);
  } else {
    static_assert(P == 4, "groups of 2 or 4 lanes");
    group_dif_stage<L, SIGN, 1, M>(v, q, (q >> 1) * M, xchg);
    group_dif_stage<L / 2, SIGN, 2, M>(v, q, 0, xchg);
  }
  fft_line<M, SIGN, T>(v);
}

}  // namespace mqpiv
