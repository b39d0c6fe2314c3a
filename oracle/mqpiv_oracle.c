/*
 * mqpiv_oracle.c -- Float64 CPU restatement of multi_quickPIV's PIV_CPU hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product path
 * is compared against, and the CPU baseline that bench.py times beside the GPU.
 * Nothing in the product library (multi_quickpiv_amd/) links, loads or calls it.
 *
 * Conventions follow the reference exactly: images are column-major with dim 1
 * (rows, "y", U) fastest; all coordinates below are 1-based like the Julia code
 * unless a name ends in 0.
 *
 * What each function restates (reference files, relative to the repo root):
 *   orc_good_size_real     src/utils/fft_ops.jl:164-198 (root, good_size_real)
 *   orc_vf_size            src/utils/piv_ops.jl:6-14    (get_vectorfield_size)
 *   orc_coords             src/utils/piv_ops.jl:38-85   (interrogation/search coords,
 *                          including the SR_TLF_off formula of :72, which places
 *                          a clipped search region 2 px early)
 *   load_windows           src/utils/piv_ops.jl:99-127, zncc_ops.jl:3-25,
 *                          masked_nsqecc_ops.jl:19-42   (padded copies)
 *   sat_build/sat_window   src/utils/common_ops.jl:150-180, 203-245
 *   similarity (FFTCC)     src/crosscorrelation_algorithms/fftcc.jl:101-126
 *              (ZNCC)      zncc.jl:117-144 + zncc_ops.jl:29-51
 *              (NSQECC)    nsqecc.jl:111-138 + nsqecc_ops.jl:1-26
 *              (masked)    mask_nsqecc.jl:126-160 + masked_nsqecc_ops.jl:123-143
 *   orc_peak               src/utils/common_ops.jl:98-124 (first_fullovp_peak)
 *   orc_subpixel           src/utils/piv_ops.jl:249-326 (gaussian_displacement)
 *   orc_sn                 src/utils/piv_ops.jl:333-367 (compute_SN; the 3-D form
 *                          throws in the reference, here it is the natural extension
 *                          [1+o-, 2SM-o+] per dim: parity unpinned in 3-D)
 *   orc_piv                src/multi_quickPIV.jl:40-96  (PIV_CPU loop), with the
 *                          masked variant following the intended semantics of
 *                          multi_quickPIV.jl:99-154 and piv_ops.jl:148-157.
 *   orc_shift_field,       multipass: the loop of multi_quickPIV.jl:57-88 with the
 *   orc_piv_multipass      scaled sizes of piv_parameters.jl:85-88 and the
 *                          search-region offset of get_search_coordinates
 *                          (piv_ops.jl:59-76), wired as piv_ops.jl:160-173
 *                          describes (the reference leaves it unfinished, so this
 *                          is the definition the device follows, stated once):
 *                          pass k = m..1 scales interSize, searchMargin, step by k;
 *                          each later pass shifts every search region by d = the
 *                          previous field interpolated multilinearly (in
 *                          vector-field index space) at the IA centre, rounded to
 *                          nearest (ties to even), NaN -> 0, clamped so that the
 *                          IA moved by d stays in the image; the search region is
 *                          the one the reference builds for an IA at TLF + d (:72
 *                          included), the IA stays at TLF, and the pass reports d
 *                          plus its displacement; SN from the last pass.
 *
 * Two correlation modes:
 *   ORC_DIRECT : C(s) = sum_x F[x] G[x+s] by direct summation in double (the exact
 *                definition; used as the parity checker).
 *   ORC_FFT    : C(s) from a complex FFT of the packed pair F + iG at the
 *                reference transform size cs = good_size_real(ss) per dim, the
 *                product conj(F^) G^ and an inverse FFT (the reference's per-IA
 *                structure: fftcc.jl:101-126; used as the timed CPU baseline).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORC_FFTCC = 0, ORC_ZNCC = 1, ORC_NSQECC = 2, ORC_MASK_NSQECC = 3 };
enum { ORC_F32 = 0, ORC_F64 = 1 };
enum { ORC_DIRECT = 0, ORC_FFT = 1 };

typedef struct {
  int32_t alg;
  int32_t ndims;
  int64_t is[3];
  int64_t sm[3];
  int64_t st[3];
  int32_t compute_sn;
  int32_t in_dtype;
  double threshold;
  double mask_threshold;
} orc_params;

/* Planted mistakes.  A test sets one (orc_set_plant) to show that the checks
 * that pin this restatement detect it, then sets ORC_PLANT_NONE again.  Each
 * is a plausible misreading of the reference:
 *   MASK_C1_UNMASKED  C1 = corr(F, G) without the mask product
 *                     (masked_nsqecc_ops.jl:19-25 multiplies F by the mask);
 *   MASK_C2_BOXSUM    C2 = the window sum of G^2 instead of corr(mask, G^2)
 *                     (mask_nsqecc.jl:155);
 *   MASK_NO_W         sum(F^2) = sum((F*m)^2) without the division by
 *                     w = m (m > 0) or 1 (mask_nsqecc.jl:144-149);
 *   ZNCC_RAW_F2       sum(F^2) of the raw IA instead of sum((F - mean)^2)
 *                     (zncc.jl:131-135 squares the mean-subtracted IA).   */
enum { ORC_PLANT_NONE = 0, ORC_PLANT_MASK_C1_UNMASKED = 1, ORC_PLANT_MASK_C2_BOXSUM = 2,
       ORC_PLANT_MASK_NO_W = 3, ORC_PLANT_ZNCC_RAW_F2 = 4 };
static int orc_plant = ORC_PLANT_NONE;
void orc_set_plant(int k) { orc_plant = k; }

typedef struct {
  int64_t tlf[3], brb[3];       /* interrogation area corners (1-based)      */
  int64_t sr_tlf[3], sr_brb[3]; /* clipped search region corners (1-based)   */
  int64_t off_lo[3], off_hi[3]; /* SR_TLF_off / SR_BRB_off                    */
} orc_coords;

/* ------------------------------------------------------------------------ */
/* fft_ops.jl:164-198                                                         */

static int64_t root_(int64_t base, int64_t goal, int64_t mul) {
  int64_t n = 0;
  int64_t p = base; /* base^(n+1) */
  while (mul * p < goal) {
    n += 1;
    p *= base;
  }
  return n;
}

int64_t orc_good_size_real(int64_t n) {
  if (n <= 6) return n;
  int64_t bestfac = 2 * n;
  int64_t lim = root_(11, bestfac, 5);
  int64_t f5 = 1;
  for (int64_t n5 = 0; n5 <= lim; ++n5, f5 *= 5) {
    int64_t x = f5;
    while (x < n) x *= 2;
    for (;;) {
      if (x < n) {
        x *= 3;
      } else if (x > n) {
        if (x < bestfac) bestfac = x;
        if (x & 1) break;
        x >>= 1;
      } else {
        return n;
      }
    }
  }
  return bestfac;
}

/* ------------------------------------------------------------------------ */
/* piv_ops.jl:6-14: length(1:st:(dim-is+1))                                   */

int64_t orc_vf_size(const orc_params* p, const int64_t* dims, int64_t* vf) {
  int64_t total = 1;
  for (int d = 0; d < p->ndims; ++d) {
    int64_t stop = dims[d] - p->is[d] + 1;
    vf[d] = (stop >= 1) ? (stop - 1) / p->st[d] + 1 : 0;
    total *= vf[d];
  }
  for (int d = p->ndims; d < 3; ++d) vf[d] = 1;
  return total;
}

/* piv_ops.jl:38-85 with common_ops.jl:54-68 (column-major linear -> cartesian) */
void orc_coords_of(const orc_params* p, const int64_t* dims, const int64_t* vf,
                   int64_t vf_idx, orc_coords* c) {
  int64_t rem = vf_idx - 1;
  for (int d = 0; d < 3; ++d) {
    c->tlf[d] = c->brb[d] = c->sr_tlf[d] = c->sr_brb[d] = 1;
    c->off_lo[d] = c->off_hi[d] = 0;
  }
  for (int d = 0; d < p->ndims; ++d) {
    int64_t v = rem % vf[d] + 1;
    rem /= vf[d];
    int64_t tlf = 1 + (v - 1) * p->st[d];
    int64_t brb = tlf + p->is[d] - 1;
    int64_t sm = p->sm[d];
    c->tlf[d] = tlf;
    c->brb[d] = brb;
    c->sr_tlf[d] = (tlf - sm > 1) ? tlf - sm : 1;
    c->sr_brb[d] = (brb + sm < dims[d]) ? brb + sm : dims[d];
    /* abs(min(0, TLF - (SM - 1)))  -- piv_ops.jl:72 (2 px short of the true count) */
    int64_t a = tlf - (sm - 1);
    c->off_lo[d] = a < 0 ? -a : 0;
    /* abs(min(0, dim - (BRB + SM))) -- piv_ops.jl:73 */
    int64_t b = dims[d] - (brb + sm);
    c->off_hi[d] = b < 0 ? -b : 0;
  }
}

/* The coordinates of IA vf_idx whose search region is shifted by d (1-based
 * TLF + d, clamped into [1, dim - is + 1]): SR corners and offsets as for an IA
 * at TLF + d; the IA corners stay.  Writes the clamped shift back into d.     */
void orc_coords_shifted(const orc_params* p, const int64_t* dims, const int64_t* vf,
                               int64_t vf_idx, int64_t* d, orc_coords* c) {
  orc_coords_of(p, dims, vf, vf_idx, c);
  for (int k = 0; k < p->ndims; ++k) {
    int64_t lo = 1 - c->tlf[k], hi = dims[k] - c->brb[k];
    if (d[k] < lo) d[k] = lo;
    if (d[k] > hi) d[k] = hi;
    int64_t tlf = c->tlf[k] + d[k], brb = c->brb[k] + d[k], sm = p->sm[k];
    c->sr_tlf[k] = (tlf - sm > 1) ? tlf - sm : 1;
    c->sr_brb[k] = (brb + sm < dims[k]) ? brb + sm : dims[k];
    int64_t a = tlf - (sm - 1);
    c->off_lo[k] = a < 0 ? -a : 0;
    int64_t b = dims[k] - (brb + sm);
    c->off_hi[k] = b < 0 ? -b : 0;
  }
}

/* ------------------------------------------------------------------------ */
/* image access                                                               */

static inline double img_at(const void* img, int dtype, int64_t lin) {
  return dtype == ORC_F32 ? (double)((const float*)img)[lin] : ((const double*)img)[lin];
}
static inline float img_at_f(const void* img, int dtype, int64_t lin) {
  return dtype == ORC_F32 ? ((const float*)img)[lin] : (float)((const double*)img)[lin];
}

/* Per-IA work buffers, sized for one IA. */
typedef struct {
  int nd;
  int64_t is[3], ss[3], ns[3]; /* ns = 2*SM+1 */
  int64_t nF, nG, nS;
  double meanF; /* ZNCC: the subtracted IA mean (planted-mistake use only) */
  double* F;   /* is^N : F, F-mean, or F*m                        */
  double* M;   /* is^N : mask IA (masked only)                     */
  double* G;   /* ss^N : G or G-mean (quirk-placed, zero padded)   */
  double* G2;  /* ss^N : G*G (masked only)                         */
  double* sat; /* (ss+1)^N                                         */
  double* S;   /* ns^N : similarity on the fully-overlapping window*/
  double* C2;  /* ns^N : second correlation (masked)               */
  /* FFT mode */
  int64_t cs[3], nC;
  double* zr;
  double* zi;
  double* tr;
  double* ti;
} orc_work;

static void work_init(orc_work* w, const orc_params* p) {
  memset(w, 0, sizeof(*w));
  w->nd = p->ndims;
  w->nF = w->nG = w->nS = w->nC = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < p->ndims) {
      w->is[d] = p->is[d];
      w->ss[d] = p->is[d] + 2 * p->sm[d];
      w->ns[d] = 2 * p->sm[d] + 1;
      int64_t cs0 = w->ss[d] > w->is[d] ? w->ss[d] : w->is[d];
      w->cs[d] = orc_good_size_real(cs0);
    } else {
      w->is[d] = w->ss[d] = w->ns[d] = w->cs[d] = 1;
    }
    w->nF *= w->is[d];
    w->nG *= w->ss[d];
    w->nS *= w->ns[d];
    w->nC *= w->cs[d];
  }
  int64_t nsat = (w->ss[0] + 1) * (w->ss[1] + 1) * (p->ndims == 3 ? w->ss[2] + 1 : 1);
  w->F = (double*)calloc(w->nF, sizeof(double));
  w->M = (double*)calloc(w->nF, sizeof(double));
  w->G = (double*)calloc(w->nG, sizeof(double));
  w->G2 = (double*)calloc(w->nG, sizeof(double));
  w->sat = (double*)calloc(nsat, sizeof(double));
  w->S = (double*)calloc(w->nS, sizeof(double));
  w->C2 = (double*)calloc(w->nS, sizeof(double));
}

static void work_init_fft(orc_work* w) {
  int64_t mx = w->cs[0];
  if (w->cs[1] > mx) mx = w->cs[1];
  if (w->cs[2] > mx) mx = w->cs[2];
  w->zr = (double*)calloc(w->nC, sizeof(double));
  w->zi = (double*)calloc(w->nC, sizeof(double));
  w->tr = (double*)calloc(4 * mx, sizeof(double));
  w->ti = (double*)calloc(4 * mx, sizeof(double));
}

static void work_free(orc_work* w) {
  free(w->F); free(w->M); free(w->G); free(w->G2); free(w->sat); free(w->S);
  free(w->C2); free(w->zr); free(w->zi); free(w->tr); free(w->ti);
}

/* ------------------------------------------------------------------------ */
/* Padded copies: piv_ops.jl:99-127, zncc_ops.jl:7-25, masked_nsqecc_ops.jl:19-42 */

static void load_windows(const orc_params* p, const int64_t* dims, const void* img1,
                         const void* img2, const void* mask, const orc_coords* c,
                         orc_work* w) {
  const int dt = p->in_dtype;
  const int64_t d0 = dims[0], d1 = p->ndims > 1 ? dims[1] : 1;
  const int zncc = (p->alg == ORC_ZNCC);
  const int masked = (p->alg == ORC_MASK_NSQECC);

  /* Interrogation window F (and mask window M). */
  double sumF = 0.0;
  if (zncc) {
    for (int64_t z = 0; z < w->is[2]; ++z)
      for (int64_t y = 0; y < w->is[1]; ++y)
        for (int64_t x = 0; x < w->is[0]; ++x) {
          int64_t lin = (c->tlf[0] - 1 + x) + d0 * ((c->tlf[1] - 1 + y) + d1 * (c->tlf[2] - 1 + z));
          sumF += img_at(img1, dt, lin);
        }
  }
  /* zncc_ops.jl:10: the mean is computed in the input element type. */
  float meanF32 = (float)sumF / (float)w->nF;
  double meanF64 = sumF / (double)w->nF;
  w->meanF = (dt == ORC_F32) ? (double)meanF32 : meanF64;
  for (int64_t z = 0; z < w->is[2]; ++z)
    for (int64_t y = 0; y < w->is[1]; ++y)
      for (int64_t x = 0; x < w->is[0]; ++x) {
        int64_t lin = (c->tlf[0] - 1 + x) + d0 * ((c->tlf[1] - 1 + y) + d1 * (c->tlf[2] - 1 + z));
        int64_t k = x + w->is[0] * (y + w->is[1] * z);
        if (zncc) {
          w->F[k] = (dt == ORC_F32) ? (double)(img_at_f(img1, dt, lin) - meanF32)
                                    : img_at(img1, dt, lin) - meanF64;
        } else if (masked) {
          double m = img_at(mask, dt, lin);
          w->M[k] = (orc_plant == ORC_PLANT_MASK_C2_BOXSUM) ? 1.0 : m;
          w->F[k] = (orc_plant == ORC_PLANT_MASK_C1_UNMASKED) ? img_at(img1, dt, lin)
                                                              : img_at(img1, dt, lin) * m;
        } else {
          w->F[k] = img_at(img1, dt, lin);
        }
      }

  /* Search window G placed at 1+SR_TLF_off (piv_ops.jl:124), zero elsewhere. */
  memset(w->G, 0, w->nG * sizeof(double));
  if (masked) memset(w->G2, 0, w->nG * sizeof(double));
  int64_t sg[3];
  for (int d = 0; d < 3; ++d) sg[d] = c->sr_brb[d] - c->sr_tlf[d] + 1;
  double sumG = 0.0;
  if (zncc) {
    for (int64_t z = 0; z < sg[2]; ++z)
      for (int64_t y = 0; y < sg[1]; ++y)
        for (int64_t x = 0; x < sg[0]; ++x) {
          int64_t lin = (c->sr_tlf[0] - 1 + x) + d0 * ((c->sr_tlf[1] - 1 + y) + d1 * (c->sr_tlf[2] - 1 + z));
          sumG += img_at(img2, dt, lin);
        }
  }
  int64_t nsg = sg[0] * sg[1] * sg[2];
  float meanG32 = (float)sumG / (float)nsg;
  double meanG64 = sumG / (double)nsg;
  for (int64_t z = 0; z < sg[2]; ++z)
    for (int64_t y = 0; y < sg[1]; ++y)
      for (int64_t x = 0; x < sg[0]; ++x) {
        int64_t lin = (c->sr_tlf[0] - 1 + x) + d0 * ((c->sr_tlf[1] - 1 + y) + d1 * (c->sr_tlf[2] - 1 + z));
        int64_t k = (x + c->off_lo[0]) +
                    w->ss[0] * ((y + (p->ndims > 1 ? c->off_lo[1] : 0)) +
                                w->ss[1] * (z + (p->ndims > 2 ? c->off_lo[2] : 0)));
        double g;
        if (zncc)
          g = (dt == ORC_F32) ? (double)(img_at_f(img2, dt, lin) - meanG32) : img_at(img2, dt, lin) - meanG64;
        else
          g = img_at(img2, dt, lin);
        w->G[k] = g;
        if (masked) w->G2[k] = g * g;
      }
}

/* Squared summed-area table of G over ss^N (common_ops.jl:150-180). */
static void sat_build(orc_work* w) {
  const int64_t h = w->ss[0], wd = w->ss[1], dp = w->ss[2];
  const int64_t H = h + 1, W = wd + 1;
  if (w->nd == 2) {
    memset(w->sat, 0, H * W * sizeof(double));
    for (int64_t cc = 1; cc <= wd; ++cc)
      for (int64_t r = 1; r <= h; ++r) {
        double v = w->G[(r - 1) + h * (cc - 1)];
        double val2 = v * v;
        w->sat[r + H * cc] = val2 + w->sat[(r - 1) + H * cc] + w->sat[r + H * (cc - 1)] -
                             w->sat[(r - 1) + H * (cc - 1)];
      }
  } else {
    const int64_t D = dp + 1;
    memset(w->sat, 0, H * W * D * sizeof(double));
#define SAT3(r, c, z) w->sat[(r) + H * ((c) + W * (z))]
    for (int64_t z = 1; z <= dp; ++z)
      for (int64_t cc = 1; cc <= wd; ++cc) {
        double tmp = 0.0;
        for (int64_t r = 1; r <= h; ++r) {
          double v = w->G[(r - 1) + h * ((cc - 1) + wd * (z - 1))];
          double val2 = v * v;
          SAT3(r, cc, z) = val2 + SAT3(r, cc - 1, z) + SAT3(r, cc, z - 1) - SAT3(r, cc - 1, z - 1) + tmp;
          tmp += val2;
        }
      }
  }
}

/* add_numerator_nsqecc (common_ops.jl:203-245): adds the window sum of G^2 for the
 * shift whose 1-based window index is (i,j,k) onto acc, in the reference's order. */
static inline double sat_window_add(const orc_work* w, double acc, int64_t i, int64_t j, int64_t k) {
  const int64_t H = w->ss[0] + 1, W = w->ss[1] + 1;
  const int64_t ih = w->is[0], iw = w->is[1];
  /* 1-based SAT index: T = i, D = i + ih; L = j, R = j + iw (Julia arrays). */
#define S2(r, c) w->sat[((r) - 1) + H * ((c) - 1)]
  if (w->nd == 2) {
    acc += S2(i + ih, j + iw);
    acc += S2(i, j);
    acc -= S2(i, j + iw);
    acc -= S2(i + ih, j);
    return acc;
  }
  const int64_t id = w->is[2];
#define S3(r, c, z) w->sat[((r) - 1) + H * (((c) - 1) + W * ((z) - 1))]
  acc += S3(i + ih, j + iw, k + id);
  acc -= S3(i, j, k);
  acc -= S3(i, j + iw, k + id);
  acc -= S3(i + ih, j, k + id);
  acc -= S3(i + ih, j + iw, k);
  acc += S3(i + ih, j, k);
  acc += S3(i, j + iw, k);
  acc += S3(i, j, k + id);
  return acc;
}

/* ------------------------------------------------------------------------ */
/* Direct correlation over the fully-overlapping window.                    */

static double corr_direct_at(const double* A, const double* B, const orc_work* w,
                             int64_t s0, int64_t s1, int64_t s2) {
  double acc = 0.0;
  for (int64_t z = 0; z < w->is[2]; ++z)
    for (int64_t y = 0; y < w->is[1]; ++y) {
      const double* a = A + w->is[0] * (y + w->is[1] * z);
      const double* b = B + s0 + w->ss[0] * ((y + s1) + w->ss[1] * (z + s2));
      for (int64_t x = 0; x < w->is[0]; ++x) acc += a[x] * b[x];
    }
  return acc;
}

static void corr_direct(const double* A, const double* B, orc_work* w, double* out) {
  for (int64_t k = 0; k < w->ns[2]; ++k)
    for (int64_t j = 0; j < w->ns[1]; ++j)
      for (int64_t i = 0; i < w->ns[0]; ++i)
        out[i + w->ns[0] * (j + w->ns[1] * k)] = corr_direct_at(A, B, w, i, j, k);
}

/* ------------------------------------------------------------------------ */
/* Mixed-radix complex FFT in double (radix 2,3,4,5 and a plain DFT for other
 * primes), Stockham out-of-place on one line.                                 */

static void fft_line(double* xr, double* xi, double* yr, double* yi, int64_t n, int sign) {
  /* Stockham autosort: ping-pong x <-> y; result ends in x. */
  double *ar = xr, *ai = xi, *br = yr, *bi = yi;
  int64_t ns = 1;
  int64_t rem = n;
  while (rem > 1) {
    int64_t r;
    if (rem % 4 == 0) r = 4;
    else if (rem % 2 == 0) r = 2;
    else if (rem % 3 == 0) r = 3;
    else if (rem % 5 == 0) r = 5;
    else {
      r = rem;
      for (int64_t f = 7; f * f <= rem; f += 2)
        if (rem % f == 0) { r = f; break; }
    }
    int64_t m = n / r;
    for (int64_t j = 0; j < m; ++j) {
      double vr[64], vi[64];
      int64_t jn = j % ns;
      double ang = sign * 2.0 * M_PI * (double)jn / (double)(ns * r);
      for (int64_t q = 0; q < r; ++q) {
        double ur = ar[j + q * m], ui = ai[j + q * m];
        double c = cos(ang * q), s = sin(ang * q);
        vr[q] = ur * c - ui * s;
        vi[q] = ur * s + ui * c;
      }
      double wr[64], wi[64];
      if (r == 2) {
        wr[0] = vr[0] + vr[1]; wi[0] = vi[0] + vi[1];
        wr[1] = vr[0] - vr[1]; wi[1] = vi[0] - vi[1];
      } else if (r == 4) {
        double ar0 = vr[0] + vr[2], ai0 = vi[0] + vi[2];
        double ar1 = vr[0] - vr[2], ai1 = vi[0] - vi[2];
        double ar2 = vr[1] + vr[3], ai2 = vi[1] + vi[3];
        double ar3 = vr[1] - vr[3], ai3 = vi[1] - vi[3];
        /* multiply (ar3, ai3) by sign*i */
        double br3 = -sign * ai3, bi3 = sign * ar3;
        wr[0] = ar0 + ar2; wi[0] = ai0 + ai2;
        wr[2] = ar0 - ar2; wi[2] = ai0 - ai2;
        wr[1] = ar1 + br3; wi[1] = ai1 + bi3;
        wr[3] = ar1 - br3; wi[3] = ai1 - bi3;
      } else {
        for (int64_t k = 0; k < r; ++k) {
          double sr = 0, si = 0;
          for (int64_t q = 0; q < r; ++q) {
            double a2 = sign * 2.0 * M_PI * (double)((k * q) % r) / (double)r;
            double c = cos(a2), s = sin(a2);
            sr += vr[q] * c - vi[q] * s;
            si += vr[q] * s + vi[q] * c;
          }
          wr[k] = sr; wi[k] = si;
        }
      }
      int64_t idx = (j / ns) * ns * r + jn;
      for (int64_t q = 0; q < r; ++q) {
        br[idx + q * ns] = wr[q];
        bi[idx + q * ns] = wi[q];
      }
    }
    double* t;
    t = ar; ar = br; br = t;
    t = ai; ai = bi; bi = t;
    ns *= r;
    rem /= r;
  }
  if (ar != xr) {
    memcpy(xr, ar, n * sizeof(double));
    memcpy(xi, ai, n * sizeof(double));
  }
}

/* N-D transform of (zr, zi) of size cs^N along each dim. */
static void fft_nd(orc_work* w, int sign) {
  const int64_t c0 = w->cs[0], c1 = w->cs[1], c2 = w->cs[2];
  double *lr = w->tr, *li = w->ti;
  int64_t mx = c0 > c1 ? c0 : c1;
  if (c2 > mx) mx = c2;
  double *yr = w->tr + 2 * mx, *yi = w->ti + 2 * mx;
  /* dim 0 (contiguous) */
  for (int64_t z = 0; z < c2; ++z)
    for (int64_t y = 0; y < c1; ++y) {
      int64_t base = c0 * (y + c1 * z);
      memcpy(lr, w->zr + base, c0 * sizeof(double));
      memcpy(li, w->zi + base, c0 * sizeof(double));
      fft_line(lr, li, yr, yi, c0, sign);
      memcpy(w->zr + base, lr, c0 * sizeof(double));
      memcpy(w->zi + base, li, c0 * sizeof(double));
    }
  if (w->nd >= 2)
    for (int64_t z = 0; z < c2; ++z)
      for (int64_t x = 0; x < c0; ++x) {
        for (int64_t y = 0; y < c1; ++y) {
          lr[y] = w->zr[x + c0 * (y + c1 * z)];
          li[y] = w->zi[x + c0 * (y + c1 * z)];
        }
        fft_line(lr, li, yr, yi, c1, sign);
        for (int64_t y = 0; y < c1; ++y) {
          w->zr[x + c0 * (y + c1 * z)] = lr[y];
          w->zi[x + c0 * (y + c1 * z)] = li[y];
        }
      }
  if (w->nd >= 3)
    for (int64_t y = 0; y < c1; ++y)
      for (int64_t x = 0; x < c0; ++x) {
        for (int64_t z = 0; z < c2; ++z) {
          lr[z] = w->zr[x + c0 * (y + c1 * z)];
          li[z] = w->zi[x + c0 * (y + c1 * z)];
        }
        fft_line(lr, li, yr, yi, c2, sign);
        for (int64_t z = 0; z < c2; ++z) {
          w->zr[x + c0 * (y + c1 * z)] = lr[z];
          w->zi[x + c0 * (y + c1 * z)] = li[z];
        }
      }
}

/* C(s) = sum_x A[x] B[x+s] via FFT at cs (fftcc.jl:101-126 / fft_ops.jl:95-105):
 * z = A + iB, Z = FFT(z); A^ = (Z(k)+conj Z(-k))/2, B^ = (Z(k)-conj Z(-k))/(2i);
 * P = conj(A^) B^; C = IFFT(P)/prod(cs).                                        */
static void corr_fft(const double* A, const double* B, orc_work* w, double* out) {
  const int64_t c0 = w->cs[0], c1 = w->cs[1], c2 = w->cs[2];
  memset(w->zr, 0, w->nC * sizeof(double));
  memset(w->zi, 0, w->nC * sizeof(double));
  for (int64_t z = 0; z < w->is[2]; ++z)
    for (int64_t y = 0; y < w->is[1]; ++y)
      for (int64_t x = 0; x < w->is[0]; ++x)
        w->zr[x + c0 * (y + c1 * z)] = A[x + w->is[0] * (y + w->is[1] * z)];
  for (int64_t z = 0; z < w->ss[2]; ++z)
    for (int64_t y = 0; y < w->ss[1]; ++y)
      for (int64_t x = 0; x < w->ss[0]; ++x)
        w->zi[x + c0 * (y + c1 * z)] = B[x + w->ss[0] * (y + w->ss[1] * z)];
  fft_nd(w, -1);
  /* pointwise product, pairing k with -k */
  for (int64_t z = 0; z < c2; ++z) {
    int64_t zn = (c2 - z) % c2;
    for (int64_t y = 0; y < c1; ++y) {
      int64_t yn = (c1 - y) % c1;
      for (int64_t x = 0; x < c0; ++x) {
        int64_t xn = (c0 - x) % c0;
        int64_t k = x + c0 * (y + c1 * z);
        int64_t kn = xn + c0 * (yn + c1 * zn);
        if (kn < k) continue;
        double zr = w->zr[k], zi = w->zi[k];
        double nr = w->zr[kn], ni = w->zi[kn];
        /* A^(k) = ((zr+nr) + i(zi-ni))/2 ; B^(k) = ((zi+ni) + i(nr-zr))/2 */
        double ar = 0.5 * (zr + nr), ai = 0.5 * (zi - ni);
        double br = 0.5 * (zi + ni), bi = 0.5 * (nr - zr);
        /* P(k) = conj(A^) * B^ */
        double pr = ar * br + ai * bi;
        double pi = ar * bi - ai * br;
        w->zr[k] = pr; w->zi[k] = pi;
        w->zr[kn] = pr; w->zi[kn] = -pi; /* P(-k) = conj(P(k)) */
      }
    }
  }
  fft_nd(w, +1);
  double scale = 1.0 / (double)w->nC;
  for (int64_t k = 0; k < w->ns[2]; ++k)
    for (int64_t j = 0; j < w->ns[1]; ++j)
      for (int64_t i = 0; i < w->ns[0]; ++i)
        out[i + w->ns[0] * (j + w->ns[1] * k)] = w->zr[i + c0 * (j + c1 * k)] * scale;
}

/* ------------------------------------------------------------------------ */
/* Similarity over the fully-overlapping window (ns^N, column-major).         */

static void similarity(const orc_params* p, orc_work* w, int mode) {
  double sumF2 = 0.0;
  if (p->alg == ORC_MASK_NSQECC) {
    /* mask_nsqecc.jl:144-149 */
    for (int64_t k = 0; k < w->nF; ++k) {
      double m = w->M[k];
      double wi = (m > 0 && orc_plant != ORC_PLANT_MASK_NO_W) ? m : 1.0;
      double fm = (orc_plant == ORC_PLANT_MASK_C1_UNMASKED) ? w->F[k] * m : w->F[k];
      sumF2 += fm * fm / wi;
    }
  } else if (p->alg != ORC_FFTCC) {
    /* zncc.jl:131-135 / nsqecc.jl:125-129 */
    for (int64_t k = 0; k < w->nF; ++k) {
      double f = w->F[k];
      if (p->alg == ORC_ZNCC && orc_plant == ORC_PLANT_ZNCC_RAW_F2) f += w->meanF;
      sumF2 += f * f;
    }
  }

  if (mode == ORC_FFT) corr_fft(w->F, w->G, w, w->S);
  else corr_direct(w->F, w->G, w, w->S);
  if (p->alg == ORC_FFTCC) return;

  if (p->alg == ORC_MASK_NSQECC) {
    if (mode == ORC_FFT) corr_fft(w->M, w->G2, w, w->C2);
    else corr_direct(w->M, w->G2, w, w->C2);
    double sq = sqrt(sumF2);
    for (int64_t k = 0; k < w->nS; ++k) {
      double v = w->S[k], c2 = w->C2[k];
      v *= -2;
      v += sumF2;
      v += c2;
      double den = sqrt(fabs(c2));
      den *= sq;
      v += den;
      w->S[k] = den / v;
    }
    return;
  }

  sat_build(w);
  double sq = sqrt(sumF2);
  for (int64_t k = 0; k < w->ns[2]; ++k)
    for (int64_t j = 0; j < w->ns[1]; ++j)
      for (int64_t i = 0; i < w->ns[0]; ++i) {
        int64_t lin = i + w->ns[0] * (j + w->ns[1] * k);
        double den = sat_window_add(w, 0.0, i + 1, j + 1, k + 1);
        den = sqrt(den);
        den *= sq;
        if (p->alg == ORC_ZNCC) {
          w->S[lin] = w->S[lin] / den;
        } else {
          double v = w->S[lin];
          v *= -2;
          v += sumF2;
          v = sat_window_add(w, v, i + 1, j + 1, k + 1);
          v += den;
          w->S[lin] = den / v;
        }
      }
}

/* ------------------------------------------------------------------------ */
/* first_fullovp_peak (common_ops.jl:98-124). Returns the peak (1-based, zeros
 * when nothing beats 0) and its value.                                       */

void orc_peak(const double* S, const int64_t* ns, int nd, const int64_t* sm,
              const int64_t* off_lo, const int64_t* off_hi, int64_t* peak, double* val) {
  int64_t lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    if (d < nd) {
      lo[d] = 1 + off_lo[d];
      hi[d] = 1 + 2 * sm[d] - off_hi[d];
    } else {
      lo[d] = hi[d] = 1;
    }
  }
  double best = 0.0;
  int64_t bp[3] = {0, 0, 0};
  for (int64_t z = lo[2]; z <= hi[2]; ++z)
    for (int64_t c = lo[1]; c <= hi[1]; ++c)
      for (int64_t r = lo[0]; r <= hi[0]; ++r) {
        double v = S[(r - 1) + ns[0] * ((c - 1) + ns[1] * (z - 1))];
        if (v > best) {
          best = v;
          bp[0] = r; bp[1] = c; bp[2] = z;
        }
      }
  for (int d = 0; d < 3; ++d) peak[d] = bp[d];
  *val = best;
}

/* Largest value in the peak window other than the peak itself (-inf if none);
 * used to classify near-ties, not part of the reference algorithm.           */
double orc_runner_up(const double* S, const int64_t* ns, int nd, const int64_t* sm,
                     const int64_t* off_lo, const int64_t* off_hi, const int64_t* peak) {
  int64_t lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = d < nd ? 1 + off_lo[d] : 1;
    hi[d] = d < nd ? 1 + 2 * sm[d] - off_hi[d] : 1;
  }
  double best = -INFINITY;
  for (int64_t z = lo[2]; z <= hi[2]; ++z)
    for (int64_t c = lo[1]; c <= hi[1]; ++c)
      for (int64_t r = lo[0]; r <= hi[0]; ++r) {
        if (r == peak[0] && c == peak[1] && (nd < 3 || z == peak[2])) continue;
        double v = S[(r - 1) + ns[0] * ((c - 1) + ns[1] * (z - 1))];
        if (v > best) best = v;
      }
  return best;
}

/* gaussian_displacement / gaussian_refinement / gaussian_2D/3D (piv_ops.jl:249-326) */
void orc_subpixel(const double* S, const int64_t* ns, int nd, const int64_t* sm,
                  const int64_t* off_lo, const int64_t* off_hi, const int64_t* peak,
                  double maxval, double* disp) {
  int interior = 1;
  for (int d = 0; d < nd; ++d) {
    if (!(peak[d] > 1 + off_lo[d])) interior = 0;
    if (!(peak[d] < 1 + 2 * sm[d] - off_hi[d])) interior = 0;
  }
#define SV(r, c, z) S[((r) - 1) + ns[0] * (((c) - 1) + ns[1] * ((z) - 1))]
  if (!interior) {
    for (int d = 0; d < nd; ++d) disp[d] = (double)peak[d] - (double)(sm[d] + 1);
    return;
  }
  int64_t r = peak[0], c = peak[1], z = nd == 3 ? peak[2] : 1;
  /* Julia `minimum` over the 3^N box; a NaN anywhere makes it NaN. */
  double mn = INFINITY;
  for (int64_t dz = (nd == 3 ? -1 : 0); dz <= (nd == 3 ? 1 : 0); ++dz)
    for (int64_t dc = -1; dc <= 1; ++dc)
      for (int64_t dr = -1; dr <= 1; ++dr) {
        double v = SV(r + dr, c + dc, z + dz);
        if (isnan(v) || isnan(mn)) mn = NAN;
        else if (v < mn) mn = v;
      }
  double lmx = log(1 + maxval - mn);
  double lm[3], lp[3];
  lm[0] = log(1 + SV(r - 1, c, z) - mn);
  lp[0] = log(1 + SV(r + 1, c, z) - mn);
  lm[1] = log(1 + SV(r, c - 1, z) - mn);
  lp[1] = log(1 + SV(r, c + 1, z) - mn);
  if (nd == 3) {
    lm[2] = log(1 + SV(r, c, z - 1) - mn);
    lp[2] = log(1 + SV(r, c, z + 1) - mn);
  }
  for (int d = 0; d < nd; ++d) {
    double g = (lm[d] - lp[d]) / (2 * lm[d] - 4 * lmx + 2 * lp[d]);
    disp[d] = ((double)peak[d] + g) - (double)(sm[d] + 1);
  }
}

/* compute_SN (piv_ops.jl:333-367): window [1+o-, 2SM-o+] per dim.             */
double orc_sn(const double* S, const int64_t* ns, int nd, const int64_t* sm,
              const int64_t* off_lo, const int64_t* off_hi) {
  int64_t lo[3], hi[3];
  int64_t len = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < nd) {
      lo[d] = 1 + off_lo[d];
      hi[d] = 2 * sm[d] - off_hi[d];
    } else {
      lo[d] = hi[d] = 1;
    }
    len *= (hi[d] >= lo[d]) ? hi[d] - lo[d] + 1 : 0;
  }
  if (len <= 0) return NAN;
  double mx = -INFINITY, sum = 0.0;
  int has_nan = 0;
  for (int64_t z = lo[2]; z <= hi[2]; ++z)
    for (int64_t c = lo[1]; c <= hi[1]; ++c)
      for (int64_t r = lo[0]; r <= hi[0]; ++r) {
        double v = S[(r - 1) + ns[0] * ((c - 1) + ns[1] * (z - 1))];
        if (isnan(v)) has_nan = 1;
        if (v > mx) mx = v;
        sum += v;
      }
  if (has_nan) return NAN;
  double avg = (sum - mx) / (double)(len - 1);
  return mx / avg;
}

/* maxval over the IA of an image (common_ops.jl:29-47) */
static double ia_maxval(const orc_params* p, const int64_t* dims, const void* img,
                        const orc_coords* c) {
  const int64_t d0 = dims[0], d1 = p->ndims > 1 ? dims[1] : 1;
  double mx = -INFINITY;
  for (int64_t z = c->tlf[2]; z <= c->brb[2]; ++z)
    for (int64_t y = c->tlf[1]; y <= c->brb[1]; ++y)
      for (int64_t x = c->tlf[0]; x <= c->brb[0]; ++x) {
        double v = img_at(img, p->in_dtype, (x - 1) + d0 * ((y - 1) + d1 * (z - 1)));
        if (v > mx) mx = v;
      }
  return mx;
}

/* skip_inter_region (piv_ops.jl:134-157) */
static int skip_ia(const orc_params* p, const int64_t* dims, const void* img1,
                   const void* mask, const orc_coords* c) {
  if (p->alg == ORC_MASK_NSQECC) {
    int skip = 0;
    if (p->threshold > 0) skip = ia_maxval(p, dims, img1, c) < p->threshold;
    if (p->mask_threshold > 0) skip = skip || ia_maxval(p, dims, mask, c) < p->mask_threshold;
    return skip;
  }
  if (p->threshold < 0) return 0;
  return ia_maxval(p, dims, img1, c) < p->threshold;
}

/* ------------------------------------------------------------------------ */
/* One IA: coordinates -> skip -> prepare -> correlate -> peak -> subpixel -> SN.
 * Returns 1 when the IA was skipped.  aux (optional, 4 doubles): peak[3], maxval. */

static int one_ia(const orc_params* p, const int64_t* dims, const int64_t* vf,
                  const void* img1, const void* img2, const void* mask,
                  const uint8_t* skip_mask, int64_t vf_idx, int mode, orc_work* w,
                  const int32_t* shift, double* disp, double* sn, double* aux) {
  orc_coords c;
  int64_t dsh[3] = {0, 0, 0};
  if (shift) {
    for (int k = 0; k < 3; ++k) dsh[k] = shift[3 * (vf_idx - 1) + k];
    orc_coords_shifted(p, dims, vf, vf_idx, dsh, &c);
  } else {
    orc_coords_of(p, dims, vf, vf_idx, &c);
  }
  int skipped = skip_mask ? (skip_mask[vf_idx - 1] != 0) : skip_ia(p, dims, img1, mask, &c);
  if (skipped) return 1;
  load_windows(p, dims, img1, img2, mask, &c, w);
  similarity(p, w, mode);
  int64_t peak[3];
  double val;
  orc_peak(w->S, w->ns, p->ndims, p->sm, c.off_lo, c.off_hi, peak, &val);
  orc_subpixel(w->S, w->ns, p->ndims, p->sm, c.off_lo, c.off_hi, peak, val, disp);
  for (int k = 0; k < p->ndims; ++k) disp[k] += (double)dsh[k];
  if (sn) *sn = orc_sn(w->S, w->ns, p->ndims, p->sm, c.off_lo, c.off_hi);
  if (aux) {
    aux[0] = (double)peak[0]; aux[1] = (double)peak[1]; aux[2] = (double)peak[2];
    aux[3] = val;
    aux[4] = orc_runner_up(w->S, w->ns, p->ndims, p->sm, c.off_lo, c.off_hi, peak);
    aux[5] = 0.0;
  }
  return 0;
}

/* PIV_CPU (multi_quickPIV.jl:40-96) over IAs [ia_begin, ia_end) (0-based),
 * writing VF (N x nVF, component fastest) and SN (nVF) as doubles.  Skipped IAs
 * are left untouched (the caller zero-initialises, as allocate_outputs does).
 * aux (optional): 6 doubles per IA (peak r,c,z, peak value, runner-up value, 0).
 * Returns 0, or -1 on bad parameters.                                          */
static int piv_shifted(const orc_params* p, const int64_t* dims, const void* img1, const void* img2,
                       const void* mask, const uint8_t* skip_mask, const int32_t* shift, int mode,
                       int nthreads, int64_t ia_begin, int64_t ia_end, double* VF, double* SN,
                       double* aux);

int orc_piv(const orc_params* p, const int64_t* dims, const void* img1, const void* img2,
            const void* mask, const uint8_t* skip_mask, int mode, int nthreads,
            int64_t ia_begin, int64_t ia_end, double* VF, double* SN, double* aux) {
  return piv_shifted(p, dims, img1, img2, mask, skip_mask, NULL, mode, nthreads, ia_begin, ia_end,
                     VF, SN, aux);
}

/* Skip flags (one byte per IA, 1 = skipped) of skip_inter_region for every IA
 * of the image (piv_ops.jl:134-157).  For the masked measure `mask` is the
 * mask the filter reads: PIV_CPU_masked with two masks passes mask_VF here
 * (multi_quickPIV.jl:186) and mask_IA to the correlation (:189), so
 * orc_piv(mask = mask_IA, skip_mask = these flags) restates that driver. */
int orc_skip_flags(const orc_params* p, const int64_t* dims, const void* img1, const void* mask,
                   uint8_t* flags) {
  if (p->ndims < 2 || p->ndims > 3) return -1;
  if (p->alg == ORC_MASK_NSQECC && p->mask_threshold > 0 && !mask) return -1;
  int64_t vf[3];
  const int64_t n = orc_vf_size(p, dims, vf);
  for (int64_t i = 0; i < n; ++i) {
    orc_coords c;
    orc_coords_of(p, dims, vf, i + 1, &c);
    flags[i] = (uint8_t)skip_ia(p, dims, img1, mask, &c);
  }
  return 0;
}

/* orc_piv with per-IA search-region shifts (3 int32 per IA; NULL = none). */
static int piv_shifted(const orc_params* p, const int64_t* dims, const void* img1, const void* img2,
                       const void* mask, const uint8_t* skip_mask, const int32_t* shift, int mode,
                       int nthreads, int64_t ia_begin, int64_t ia_end, double* VF, double* SN,
                       double* aux) {
  if (p->ndims < 2 || p->ndims > 3) return -1;
  if (p->alg == ORC_MASK_NSQECC && !mask) return -1;
  int64_t vf[3];
  int64_t nvf = orc_vf_size(p, dims, vf);
  if (ia_end < 0 || ia_end > nvf) ia_end = nvf;
  if (ia_begin < 0) ia_begin = 0;
  const int nd = p->ndims;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
  {
    orc_work w;
    work_init(&w, p);
    if (mode == ORC_FFT) work_init_fft(&w);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (int64_t i = ia_begin; i < ia_end; ++i) {
      double disp[3] = {0, 0, 0}, sn = 0.0;
      int skipped = one_ia(p, dims, vf, img1, img2, mask, skip_mask, i + 1, mode, &w, shift, disp,
                           SN ? &sn : NULL, aux ? aux + 6 * i : NULL);
      if (skipped) continue;
      for (int d = 0; d < nd; ++d) VF[d + nd * i] = disp[d];
      if (SN) SN[i] = sn;
    }
    work_free(&w);
  }
  return 0;
}

/* The similarity window of one IA (ns^N doubles, column-major) plus its
 * coordinates; used by tests to pin the correlation arithmetic.             */
int orc_ia_similarity(const orc_params* p, const int64_t* dims, const void* img1,
                      const void* img2, const void* mask, int64_t vf_idx, int mode,
                      double* out, int64_t* coords18) {
  int64_t vf[3];
  orc_vf_size(p, dims, vf);
  orc_coords c;
  orc_coords_of(p, dims, vf, vf_idx, &c);
  orc_work w;
  work_init(&w, p);
  if (mode == ORC_FFT) work_init_fft(&w);
  load_windows(p, dims, img1, img2, mask, &c, &w);
  similarity(p, &w, mode);
  memcpy(out, w.S, w.nS * sizeof(double));
  if (coords18) {
    for (int d = 0; d < 3; ++d) {
      coords18[d] = c.tlf[d]; coords18[3 + d] = c.brb[d];
      coords18[6 + d] = c.sr_tlf[d]; coords18[9 + d] = c.sr_brb[d];
      coords18[12 + d] = c.off_lo[d]; coords18[15 + d] = c.off_hi[d];
    }
  }
  work_free(&w);
  return 0;
}

/* Raw correlation C(s) of one IA (no normalisation), for tests. */
int orc_ia_correlation(const orc_params* p, const int64_t* dims, const void* img1,
                       const void* img2, int64_t vf_idx, int mode, double* out) {
  orc_params q = *p;
  q.alg = ORC_FFTCC;
  return orc_ia_similarity(&q, dims, img1, img2, NULL, vf_idx, mode, out, NULL);
}

/* ------------------------------------------------------------------------ */
/* Multipass (definition in the header of this file).                        */

/* Search-region shifts (3 int32 per fine IA, 0 beyond nd) from the coarse
 * field vfc (nd x nVF_c doubles, component fastest).  Each dim places the fine
 * IA's centre v*stf + (isf-1)/2 (0-based pixels) in coarse index units:
 * t = (2*v*stf + (isf-1) - (isc-1)) / (2*stc), clamped to [0, nc-1]; corners
 * a = min(floor(t), max(nc-2, 0)), b = min(a+1, nc-1), weight of b = t - a.
 * The corners are summed with b0 fastest, weights multiplied dim 0 first.   */
void orc_shift_field(int nd, const double* vfc, const int64_t* nc, const int64_t* isc,
                     const int64_t* stc, const int64_t* nf, const int64_t* isf, const int64_t* stf,
                     int32_t* out) {
  int64_t n = 1;
  for (int k = 0; k < nd; ++k) n *= nf[k];
  for (int64_t i = 0; i < n; ++i) {
    int64_t ia[3] = {0, 0, 0}, ib[3] = {0, 0, 0};
    double w[3] = {0.0, 0.0, 0.0};
    int64_t r = i;
    for (int k = 0; k < nd; ++k) {
      int64_t v = r % nf[k];
      r /= nf[k];
      int64_t num = 2 * v * stf[k] + (isf[k] - 1) - (isc[k] - 1);
      double t = (double)num / (double)(2 * stc[k]);
      double tmax = (double)(nc[k] - 1);
      t = t < 0.0 ? 0.0 : (t > tmax ? tmax : t);
      int64_t a = (int64_t)floor(t);
      int64_t top = nc[k] >= 2 ? nc[k] - 2 : 0;
      if (a > top) a = top;
      ia[k] = a;
      ib[k] = a + 1 < nc[k] ? a + 1 : nc[k] - 1;
      w[k] = t - (double)a;
    }
    for (int c = 0; c < 3; ++c) {
      int32_t d = 0;
      if (c < nd) {
        double acc = 0.0;
        for (int b2 = 0; b2 < (nd == 3 ? 2 : 1); ++b2)
          for (int b1 = 0; b1 < 2; ++b1)
            for (int b0 = 0; b0 < 2; ++b0) {
              double wt = (b0 ? w[0] : 1.0 - w[0]) * (b1 ? w[1] : 1.0 - w[1]);
              if (nd == 3) wt = wt * (b2 ? w[2] : 1.0 - w[2]);
              int64_t u = (b0 ? ib[0] : ia[0]) + nc[0] * ((b1 ? ib[1] : ia[1]) + nc[1] * (b2 ? ib[2] : ia[2]));
              acc = acc + wt * vfc[nd * u + c];
            }
        if (acc == acc) {
          const double lim = 1073741824.0;
          acc = acc > lim ? lim : (acc < -lim ? -lim : acc);
          d = (int32_t)nearbyint(acc);
        }
      }
      out[3 * i + c] = d;
    }
  }
}

/* Multipass PIV over all IAs: VF/SN of the last pass (the caller zero-fills
 * them).  Returns 0, or -1 on bad parameters.                               */
int orc_piv_multipass(const orc_params* p, int m, const int64_t* dims, const void* img1,
                      const void* img2, const void* mask, int mode, int nthreads, double* VF,
                      double* SN) {
  if (m < 1 || p->ndims < 2 || p->ndims > 3) return -1;
  const int nd = p->ndims;
  double* prev = NULL;
  int64_t pn[3] = {1, 1, 1}, pis[3] = {1, 1, 1}, pst[3] = {1, 1, 1};
  for (int k = m; k >= 1; --k) {
    orc_params q = *p;
    for (int d = 0; d < nd; ++d) {
      q.is[d] *= k;
      q.sm[d] *= k;
      q.st[d] *= k;
    }
    if (k > 1) q.compute_sn = 0;
    int64_t vf[3];
    int64_t n = orc_vf_size(&q, dims, vf);
    int32_t* shift = NULL;
    if (prev && n > 0) {
      shift = (int32_t*)calloc(3 * n, sizeof(int32_t));
      orc_shift_field(nd, prev, pn, pis, pst, vf, q.is, q.st, shift);
    }
    double* out = VF;
    double* sn = SN;
    if (k > 1) {
      out = (double*)calloc(n * nd > 0 ? n * nd : 1, sizeof(double));
      sn = NULL;
    }
    int rc = n > 0 ? piv_shifted(&q, dims, img1, img2, mask, NULL, shift, mode, nthreads, 0, n, out,
                                 q.compute_sn ? sn : NULL, NULL)
                   : 0;
    free(shift);
    free(prev);
    prev = NULL;
    if (rc) {
      if (k > 1) free(out);
      return rc;
    }
    if (k > 1) {
      if (n > 0) {
        prev = out;
        for (int d = 0; d < 3; ++d) {
          pn[d] = vf[d];
          pis[d] = q.is[d];
          pst[d] = q.st[d];
        }
      } else {
        free(out);
      }
    }
  }
  free(prev);
  return 0;
}
