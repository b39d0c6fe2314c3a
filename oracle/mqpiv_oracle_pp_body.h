/*
 * mqpiv_oracle_pp_body.h -- element-type-generic body of the post-processing
 * oracle, included by mqpiv_oracle_pp.c once per element type with
 *   PP_T    the vector-field element type (float or double),
 *   PP_SFX  the function-name suffix (_f32 / _f64).
 * TEST INFRASTRUCTURE ONLY (see mqpiv_oracle_pp.c).
 */
#define PP_CAT2(a, b) a##b
#define PP_CAT(a, b) PP_CAT2(a, b)
#define PP_FN(name) PP_CAT(name, PP_SFX)

/* magnitudes(VF): sqrt(sum(VF[:,i].^2)) in the element type, components summed
 * in order (postprocessing.jl:178-194). */
void PP_FN(orc_pp_magnitudes)(int nd, const int64_t* sz, const PP_T* VF, PP_T* M) {
  const int64_t nv = pp_count(nd, sz);
  for (int64_t v = 0; v < nv; ++v) M[v] = pp_mag_t(VF + v * nd, nd);
}

/* TH_masking(th, map; cmp) (postprocessing.jl:640-663): cmp 0 is x > th (the
 * default), 1 is x < th (remove_low_SN, :712-732). */
void PP_FN(orc_pp_th_mask)(int64_t n, const PP_T* x, double th, int cmp, uint8_t* mask) {
  for (int64_t i = 0; i < n; ++i) mask[i] = cmp ? ((double)x[i] < th) : ((double)x[i] > th);
}

/* STD_masking(n, map) (postprocessing.jl:598-610): |x - mean| > n * std with the
 * corrected (n - 1) standard deviation.  Sums in double, sequential. */
void PP_FN(orc_pp_std_mask)(int64_t n, const PP_T* x, double nsd, uint8_t* mask, double* stats) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += (double)x[i];
  const double mean = s / (double)n;
  double q = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double d = (double)x[i] - mean;
    q += d * d;
  }
  const double sd = sqrt(q / (double)(n - 1));
  for (int64_t i = 0; i < n; ++i) mask[i] = fabs((double)x[i] - mean) > nsd * sd;
  if (stats) {
    stats[0] = mean;
    stats[1] = sd;
  }
}

/* STD_filtering / TH_filtering / remove_low_SN replacement step
 * (postprocessing.jl:539-573, 618-634, 672-703): flagged vectors get the zero
 * vector (how 0, zeroReplace) or the per-component median over the clipped box
 * of radius r around them in the ORIGINAL field (how 1: medianReplace, and
 * meanReplace, whose body is the median too). */
void PP_FN(orc_pp_replace)(int nd, const int64_t* sz, const PP_T* VF, const uint8_t* mask, int how,
                           const int64_t* r, PP_T* out) {
  const int64_t nv = pp_count(nd, sz);
  int64_t cap = 1;
  for (int d = 0; d < nd; ++d) cap *= 2 * r[d] + 1;
  PP_T* buf = (PP_T*)malloc(sizeof(PP_T) * (size_t)cap);
  for (int64_t v = 0; v < nv; ++v) {
    if (!mask[v]) {
      for (int c = 0; c < nd; ++c) out[v * nd + c] = VF[v * nd + c];
      continue;
    }
    if (how == 0) {
      for (int c = 0; c < nd; ++c) out[v * nd + c] = (PP_T)0;
      continue;
    }
    int64_t lo[3], hi[3];
    pp_box(nd, sz, v, r, lo, hi);
    for (int c = 0; c < nd; ++c) {
      int64_t m = 0;
      for (int64_t z = lo[2]; z <= hi[2]; ++z)
        for (int64_t y = lo[1]; y <= hi[1]; ++y)
          for (int64_t x = lo[0]; x <= hi[0]; ++x) buf[m++] = VF[(x + sz[0] * (y + sz[1] * z)) * nd + c];
      out[v * nd + c] = PP_FN(pp_median)(buf, m);
    }
  }
  free(buf);
}

/* average(radius, VF) (postprocessing.jl:359-381): per-component mean over the
 * clipped box, summed in the element type in column-major order, divided by the
 * count. */
void PP_FN(orc_pp_average)(int nd, const int64_t* sz, const PP_T* VF, const int64_t* r, PP_T* out) {
  const int64_t nv = pp_count(nd, sz);
  for (int64_t v = 0; v < nv; ++v) {
    int64_t lo[3], hi[3];
    pp_box(nd, sz, v, r, lo, hi);
    for (int c = 0; c < nd; ++c) {
      PP_T s = 0;
      int64_t m = 0;
      for (int64_t z = lo[2]; z <= hi[2]; ++z)
        for (int64_t y = lo[1]; y <= hi[1]; ++y)
          for (int64_t x = lo[0]; x <= hi[0]; ++x) {
            const PP_T a = VF[(x + sz[0] * (y + sz[1] * z)) * nd + c];
            s = m ? (PP_T)(s + a) : a;
            ++m;
          }
      out[v * nd + c] = (PP_T)(s / (PP_T)m);
    }
  }
}

/* Normalised dot product of the unit vectors at v and u, in the element type:
 * sum over c of (VF[c,v]/M[v]) * (VF[c,u]/M[u]), products rounded, summed in
 * order (postprocessing.jl:316-322). */
static PP_T PP_FN(pp_ndot)(const PP_T* a, PP_T ma, const PP_T* b, PP_T mb, int nd) {
  PP_T s = (PP_T)((PP_T)(a[0] / ma) * (PP_T)(b[0] / mb));
  for (int c = 1; c < nd; ++c) s = (PP_T)(s + (PP_T)((PP_T)(a[c] / ma) * (PP_T)(b[c] / mb)));
  return s;
}

/* collectiveness(radius, VF; max_angle) (postprocessing.jl:294-331): fraction of
 * the clipped box (the vector itself included) whose normalised dot product
 * with the vector exceeds min_dot = cosd(max_angle); Float32 output.  Zero
 * vectors normalise to NaN and never count. */
void PP_FN(orc_pp_collectiveness)(int nd, const int64_t* sz, const PP_T* VF, const int64_t* r, double min_dot,
                                  float* out) {
  const int64_t nv = pp_count(nd, sz);
  PP_T* M = (PP_T*)malloc(sizeof(PP_T) * (size_t)nv);
  PP_FN(orc_pp_magnitudes)(nd, sz, VF, M);
  for (int64_t v = 0; v < nv; ++v) {
    int64_t lo[3], hi[3];
    pp_box(nd, sz, v, r, lo, hi);
    int64_t nn = 1, ns = 0;
    for (int d = 0; d < 3; ++d) nn *= hi[d] - lo[d] + 1;
    for (int64_t z = lo[2]; z <= hi[2]; ++z)
      for (int64_t y = lo[1]; y <= hi[1]; ++y)
        for (int64_t x = lo[0]; x <= hi[0]; ++x) {
          const int64_t u = x + sz[0] * (y + sz[1] * z);
          ns += (double)PP_FN(pp_ndot)(VF + v * nd, M[v], VF + u * nd, M[u], nd) > min_dot;
        }
    out[v] = (float)((double)ns / (double)nn);
  }
  free(M);
}

/* similarityAveraging(radius, VF; max_angle, normalize, scale_by_similarity,
 * similarity_exponent) (postprocessing.jl:478-523): for each non-zero vector,
 * the sum over similar neighbours of (VF[:,u]/M[u])*M[u] (element type,
 * accumulated in double), times 1/|sum| when normalising and times
 * (N_similar/N_neighbors)^exponent when scaling.  Zero vectors are copied. */
void PP_FN(orc_pp_simavg)(int nd, const int64_t* sz, const PP_T* VF, const int64_t* r, double min_dot,
                          int normalize, int scale, int expo, PP_T* out) {
  const int64_t nv = pp_count(nd, sz);
  PP_T* M = (PP_T*)malloc(sizeof(PP_T) * (size_t)nv);
  PP_FN(orc_pp_magnitudes)(nd, sz, VF, M);
  for (int64_t v = 0; v < nv; ++v) {
    if (M[v] == 0) {
      for (int c = 0; c < nd; ++c) out[v * nd + c] = VF[v * nd + c];
      continue;
    }
    int64_t lo[3], hi[3];
    pp_box(nd, sz, v, r, lo, hi);
    int64_t nn = 1, ns = 0;
    for (int d = 0; d < 3; ++d) nn *= hi[d] - lo[d] + 1;
    double avg[3] = {0.0, 0.0, 0.0};
    for (int64_t z = lo[2]; z <= hi[2]; ++z)
      for (int64_t y = lo[1]; y <= hi[1]; ++y)
        for (int64_t x = lo[0]; x <= hi[0]; ++x) {
          const int64_t u = x + sz[0] * (y + sz[1] * z);
          const PP_T dot = PP_FN(pp_ndot)(VF + v * nd, M[v], VF + u * nd, M[u], nd);
          if ((double)dot > min_dot) {
            for (int c = 0; c < nd; ++c) avg[c] += (double)(PP_T)((PP_T)(VF[u * nd + c] / M[u]) * M[u]);
            ++ns;
          }
        }
    double fac = 1.0;
    if (normalize) {
      double q = avg[0] * avg[0];
      for (int c = 1; c < nd; ++c) q += avg[c] * avg[c];
      fac = 1.0 / sqrt(q);
    }
    if (scale) fac *= pp_powi((double)ns / (double)nn, expo);
    for (int c = 0; c < nd; ++c) out[v * nd + c] = (PP_T)(avg[c] * fac);
  }
  free(M);
}

/* local_magnitude_mask(ratio, VF; radius) (postprocessing.jl:202-218): keep a
 * vector iff M / mean(M over the clipped box without itself) < ratio; dropped
 * vectors become copysign(0, v) (Float * false); removed = !keep.  The
 * reference indexes VF_c[c,:,:] and so throws for 3-D fields; the 3-D form here
 * is the natural extension (parity unpinned). */
void PP_FN(orc_pp_lmm)(int nd, const int64_t* sz, const PP_T* VF, double ratio, const int64_t* r, PP_T* out,
                       uint8_t* removed) {
  const int64_t nv = pp_count(nd, sz);
  PP_T* M = (PP_T*)malloc(sizeof(PP_T) * (size_t)nv);
  PP_FN(orc_pp_magnitudes)(nd, sz, VF, M);
  for (int64_t v = 0; v < nv; ++v) {
    int64_t lo[3], hi[3];
    pp_box(nd, sz, v, r, lo, hi);
    PP_T s = 0;
    int64_t m = 0;
    for (int64_t z = lo[2]; z <= hi[2]; ++z)
      for (int64_t y = lo[1]; y <= hi[1]; ++y)
        for (int64_t x = lo[0]; x <= hi[0]; ++x) {
          const PP_T a = M[x + sz[0] * (y + sz[1] * z)];
          s = m ? (PP_T)(s + a) : a;
          ++m;
        }
    const PP_T sum_mag = (PP_T)(s - M[v]);
    const PP_T mean_mag = (PP_T)(sum_mag / (PP_T)(m - 1));
    const int keep = (double)(PP_T)(M[v] / mean_mag) < ratio;
    for (int c = 0; c < nd; ++c) {
      const PP_T a = VF[v * nd + c];
      out[v * nd + c] = keep ? a : (PP_T)copysign(0.0, (double)a);
    }
    removed[v] = !keep;
  }
  free(M);
}

#undef PP_FN
#undef PP_CAT
#undef PP_CAT2
