/*
 * mqpiv_oracle_pp.c -- CPU restatement of multi_quickPIV's vector-field
 * post-processing filters (src/postprocessing/postprocessing.jl).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP filters in
 * multi_quickpiv_amd/ are compared against; the product library never links,
 * loads or calls it.
 *
 * Layout follows the reference: VF is (N, nVF_1, nVF_2[, nVF_3]) column-major,
 * component fastest; per-vector maps (magnitudes, SN, masks) are nVF... with
 * dim 1 fastest.  Boxes are the reference's clipped ROI
 * UnitRange.(max.(1, c .- r), min.(size, c .+ r)) (postprocessing.jl:63-70).
 *
 * Element-type arithmetic follows the Julia code for Float64 and Float32
 * fields; where Julia's Base.sum / Statistics reduce in an order that depends
 * on SIMD width (sum with @simd, pairwise blocks), this restatement sums
 * sequentially in column-major order -- equal to the reference up to rounding
 * (parity unpinned at the last-bit level for sums; exact everywhere else).
 *
 * Reference behaviours that are not reproduced because they raise in Julia:
 *   - meanReplace(coords, radius, VF) splats VF into a tuple and calls localOp
 *     with it, which has no method: STD_filtering(VF) with its default
 *     replaceFun throws once a vector is flagged.  Here meanReplace follows its
 *     body (the median), as medianReplace does.
 *   - TH_filtering calls the undefined `replacefun`; remove_low_SN passes the
 *     unsupported keyword `replace`.  Here both follow their intended semantics
 *     (mask, then the chosen replacement).
 *   - local_magnitude_mask indexes VF_c[c,:,:] and throws for 3-D fields; the
 *     3-D form here is the natural extension.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int64_t pp_count(int nd, const int64_t* sz) {
  int64_t n = 1;
  for (int d = 0; d < nd; ++d) n *= sz[d];
  return n;
}

/* 0-based inclusive clipped box of radius r around vector v; unused dims 0..0. */
static void pp_box(int nd, const int64_t* sz, int64_t v, const int64_t* r, int64_t* lo, int64_t* hi) {
  for (int d = 0; d < 3; ++d) {
    lo[d] = hi[d] = 0;
    if (d >= nd) continue;
    const int64_t c = v % sz[d];
    v /= sz[d];
    lo[d] = c - r[d] < 0 ? 0 : c - r[d];
    hi[d] = c + r[d] > sz[d] - 1 ? sz[d] - 1 : c + r[d];
  }
}

/* x^n for integer n >= 0 as Julia's ^(::Float64, ::Integer): n <= 3 by plain
 * products (x*x, x*x*x), larger n by compensated power-by-squaring with
 * error-free products (Graillat, Langlois & Louvet, "Accurate floating point
 * product and exponentiation", 2009).  Parity unpinned for n >= 4. */
static double pp_powi(double x, int n) {
  if (n == 0) return 1.0;
  if (n == 1) return x;
  if (n == 2) return x * x;
  if (n == 3) return x * x * x;
  double y = 1.0, xlo = 0.0, ylo = 0.0;
  while (n > 1) {
    if (n & 1) {
      const double err = fma(y, xlo, x * ylo);
      const double p = x * y;
      ylo = fma(x, y, -p) + err;
      y = p;
    }
    const double err = x * 2 * xlo;
    const double p = x * x;
    xlo = fma(x, x, -p) + err;
    x = p;
    n >>= 1;
  }
  const double err = fma(y, xlo, x * ylo);
  return (isfinite(x) && isfinite(err)) ? fma(x, y, err) : x * y;
}

/* Total order of Base.isless on non-NaN floats: -0.0 sorts before +0.0. */
static int pp_cmp_d(const void* a, const void* b) {
  const double x = *(const double*)a, y = *(const double*)b;
  if (x < y) return -1;
  if (x > y) return 1;
  return (int)(signbit(y) != 0) - (int)(signbit(x) != 0);
}
static int pp_cmp_f(const void* a, const void* b) {
  const float x = *(const float*)a, y = *(const float*)b;
  if (x < y) return -1;
  if (x > y) return 1;
  return (int)(signbit(y) != 0) - (int)(signbit(x) != 0);
}

/* Statistics.median: NaN if any value is NaN; odd count: the middle element;
 * even: middle(a, b) = a/2 + b/2 of the two central order statistics. */
static double pp_median_f64(double* v, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (isnan(v[i])) return NAN;
  qsort(v, (size_t)n, sizeof(double), pp_cmp_d);
  if (n & 1) return v[n / 2];
  return v[n / 2 - 1] / 2 + v[n / 2] / 2;
}
static float pp_median_f32(float* v, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (isnan(v[i])) return NAN;
  qsort(v, (size_t)n, sizeof(float), pp_cmp_f);
  if (n & 1) return v[n / 2];
  return v[n / 2 - 1] / 2.0f + v[n / 2] / 2.0f;
}

static double pp_mag_t_f64(const double* p, int nd) {
  double s = p[0] * p[0];
  for (int c = 1; c < nd; ++c) s = s + p[c] * p[c];
  return sqrt(s);
}
static float pp_mag_t_f32(const float* p, int nd) {
  float s = p[0] * p[0];
  for (int c = 1; c < nd; ++c) s = s + p[c] * p[c];
  return sqrtf(s);
}

#define PP_T double
#define PP_SFX _f64
#define pp_mag_t pp_mag_t_f64
#include "mqpiv_oracle_pp_body.h"
#undef pp_mag_t
#undef PP_T
#undef PP_SFX

#define PP_T float
#define PP_SFX _f32
#define pp_mag_t pp_mag_t_f32
#include "mqpiv_oracle_pp_body.h"
#undef pp_mag_t
#undef PP_T
#undef PP_SFX
