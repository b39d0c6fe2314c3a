// mqpiv_post_api.cpp -- C ABI of the device vector-field filters
// (src/postprocessing/postprocessing.jl).  Each entry point validates the
// field description, takes its scratch from the context and enqueues the
// kernels of mqpiv_post.hip on the given stream without synchronising.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "../../include/mqpiv.h"
#include "mqpiv_ctx.h"
#include "mqpiv_post.h"

using mqpiv::post::VGrid;

namespace {

// Validates desc (and radius when given) and fills the launch grid.
int make_grid(mqpiv_ctx* c, const mqpiv_vf_desc* d, const int64_t* radius, VGrid* g) {
  if (!d) return fail(c, MQPIV_ERR_NULL, "field description is NULL");
  if (d->ndims != 2 && d->ndims != 3) return fail(c, MQPIV_ERR_BAD_PARAMS, "ndims must be 2 or 3");
  if (d->dtype != MQPIV_F32 && d->dtype != MQPIV_F64)
    return fail(c, MQPIV_ERR_BAD_PARAMS, "dtype must be MQPIV_F32 or MQPIV_F64");
  g->nd = d->ndims;
  g->nv = 1;
  for (int k = 0; k < 3; ++k) {
    const int64_t n = k < d->ndims ? d->size[k] : 1;
    if (n < 0 || n > (int64_t)INT32_MAX) return fail(c, MQPIV_ERR_BAD_PARAMS, "size[%d] out of range", k);
    g->n[k] = (int32_t)n;
    g->nv *= n;
    int64_t r = 0;
    if (radius && k < d->ndims) r = radius[k];
    if (r < 0 || r > 1024) return fail(c, MQPIV_ERR_BAD_PARAMS, "radius[%d] must be in [0, 1024]", k);
    g->r[k] = (int32_t)r;
  }
  return MQPIV_OK;
}

size_t esize(const mqpiv_vf_desc* d) { return d->dtype == MQPIV_F32 ? 4 : 8; }

// Context scratch: [reduction partials | magnitude map of nv elements].
int pp_scratch(mqpiv_ctx* c, const mqpiv_vf_desc* d, int64_t nv, hipStream_t s, double** red, void** M) {
  const size_t rb = (mqpiv::pp_std_scratch_bytes() + 255) & ~(size_t)255;
  int rc = ensure(c, &c->d_pp, &c->pp_cap, rb + (size_t)nv * esize(d) + 256, s);
  if (rc) return rc;
  *red = (double*)c->d_pp;
  *M = (char*)c->d_pp + rb;
  return MQPIV_OK;
}

int done(mqpiv_ctx* c, hipError_t e, const char* what) {
  if (e != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
  return MQPIV_OK;
}

#define PP_ENTER(c, stream)                                                          \
  if (!c) return MQPIV_ERR_NULL;                                                     \
  std::lock_guard<std::mutex> lk(c->mu);                                             \
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "hipSetDevice failed"); \
  hipStream_t s = stream ? (hipStream_t)stream : c->stream

}  // namespace

extern "C" int mqpiv_pp_magnitudes(mqpiv_ctx* c, const mqpiv_vf_desc* d, const void* VF, void* M, void* stream) {
  PP_ENTER(c, stream);
  VGrid g;
  int rc = make_grid(c, d, nullptr, &g);
  if (rc) return rc;
  if (g.nv && (!VF || !M)) return fail(c, MQPIV_ERR_NULL, "VF or M is NULL");
  return done(c, mqpiv::pp_magnitudes(d->dtype, g, VF, M, s), "magnitudes");
}

extern "C" int mqpiv_pp_std_mask(mqpiv_ctx* c, const mqpiv_vf_desc* d, const void* map, double n_std,
                                 uint8_t* mask, void* stream) {
  PP_ENTER(c, stream);
  VGrid g;
  int rc = make_grid(c, d, nullptr, &g);
  if (rc) return rc;
  if (g.nv && (!map || !mask)) return fail(c, MQPIV_ERR_NULL, "map or mask is NULL");
  double* red;
  void* M;
  rc = pp_scratch(c, d, 0, s, &red, &M);
  if (rc) return rc;
  return done(c, mqpiv::pp_std_mask(d->dtype, g.nv, map, n_std, mask, red, s), "STD mask");
}

extern "C" int mqpiv_pp_th_mask(mqpiv_ctx* c, const mqpiv_vf_desc* d, const void* map, double th, int32_t cmp,
                                uint8_t* mask, void* stream) {
  PP_ENTER(c, stream);
  VGrid g;
  int rc = make_grid(c, d, nullptr, &g);
  if (rc) return rc;
  if (cmp != MQPIV_CMP_GT && cmp != MQPIV_CMP_LT) return fail(c, MQPIV_ERR_BAD_PARAMS, "cmp must be GT or LT");
  if (g.nv && (!map || !mask)) return fail(c, MQPIV_ERR_NULL, "map or mask is NULL");
  return done(c, mqpiv::pp_th_mask(d->dtype, g.nv, map, th, cmp, mask, s), "TH mask");
}

extern "C" int mqpiv_pp_replace(mqpiv_ctx* c, const mqpiv_vf_desc* d, const void* VF, const uint8_t* mask,
                                int32_t how, const int64_t* radius, void* out, void* stream) {
  PP_ENTER(c, stream);
  VGrid g;
  int rc = make_grid(c, d, radius, &g);
  if (rc) return rc;
  if (how != MQPIV_REPLACE_ZERO && how != MQPIV_REPLACE_MEDIAN)
    return fail(c, MQPIV_ERR_BAD_PARAMS, "how must be MQPIV_REPLACE_ZERO or MQPIV_REPLACE_MEDIAN");
  if (how == MQPIV_REPLACE_MEDIAN && !radius) return fail(c, MQPIV_ERR_NULL, "radius is NULL");
  if (g.nv && (!VF || !mask || !out)) return fail(c, MQPIV_ERR_NULL, "VF, mask or out is NULL");
  if (g.nv && out == VF) return fail(c, MQPIV_ERR_BAD_PARAMS, "out must not alias VF");
  return done(c, mqpiv::pp_replace(d->dtype, g, VF, mask, how, out, s), "replace");
}

extern "C" int mqpiv_pp_average(mqpiv_ctx* c, const mqpiv_vf_desc* d, const void* VF, const int64_t* radius,
                                void* out, void* stream) {
  PP_ENTER(c, stream);
  if (!radius) return fail(c, MQPIV_ERR_NULL, "radius is NULL");
  VGrid g;
  int rc = make_grid(c, d, radius, &g);
  if (rc) return rc;
  if (g.nv && (!VF || !out)) return fail(c, MQPIV_ERR_NULL, "VF or out is NULL");
  if (g.nv && out == VF) return fail(c, MQPIV_ERR_BAD_PARAMS, "out must not alias VF");
  return done(c, mqpiv::pp_average(d->dtype, g, VF, out, s), "average");
}

extern "C" int mqpiv_pp_collectiveness(mqpiv_ctx* c, const mqpiv_vf_desc* d, const void* VF,
                                       const int64_t* radius, double min_dot, float* out, void* stream) {
  PP_ENTER(c, stream);
  if (!radius) return fail(c, MQPIV_ERR_NULL, "radius is NULL");
  VGrid g;
  int rc = make_grid(c, d, radius, &g);
  if (rc) return rc;
  if (g.nv && (!VF || !out)) return fail(c, MQPIV_ERR_NULL, "VF or out is NULL");
  double* red;
  void* M;
  rc = pp_scratch(c, d, g.nv, s, &red, &M);
  if (rc) return rc;
  return done(c, mqpiv::pp_collectiveness(d->dtype, g, VF, M, min_dot, out, s), "collectiveness");
}

extern "C" int mqpiv_pp_similarity_averaging(mqpiv_ctx* c, const mqpiv_vf_desc* d, const void* VF,
                                             const int64_t* radius, double min_dot, int32_t normalize,
                                             int32_t scale_by_similarity, int32_t similarity_exponent, void* out,
                                             void* stream) {
  PP_ENTER(c, stream);
  if (!radius) return fail(c, MQPIV_ERR_NULL, "radius is NULL");
  VGrid g;
  int rc = make_grid(c, d, radius, &g);
  if (rc) return rc;
  if (similarity_exponent < 0) return fail(c, MQPIV_ERR_UNSUPPORTED, "negative similarity_exponent");
  if (g.nv && (!VF || !out)) return fail(c, MQPIV_ERR_NULL, "VF or out is NULL");
  if (g.nv && out == VF) return fail(c, MQPIV_ERR_BAD_PARAMS, "out must not alias VF");
  double* red;
  void* M;
  rc = pp_scratch(c, d, g.nv, s, &red, &M);
  if (rc) return rc;
  return done(c,
              mqpiv::pp_simavg(d->dtype, g, VF, M, min_dot, normalize ? 1 : 0, scale_by_similarity ? 1 : 0,
                               similarity_exponent, out, s),
              "similarity averaging");
}

extern "C" int mqpiv_pp_local_magnitude_mask(mqpiv_ctx* c, const mqpiv_vf_desc* d, const void* VF, double ratio,
                                             const int64_t* radius, void* out, uint8_t* removed, void* stream) {
  PP_ENTER(c, stream);
  if (!radius) return fail(c, MQPIV_ERR_NULL, "radius is NULL");
  VGrid g;
  int rc = make_grid(c, d, radius, &g);
  if (rc) return rc;
  if (g.nv && (!VF || !out || !removed)) return fail(c, MQPIV_ERR_NULL, "VF, out or removed is NULL");
  if (g.nv && out == VF) return fail(c, MQPIV_ERR_BAD_PARAMS, "out must not alias VF");
  double* red;
  void* M;
  rc = pp_scratch(c, d, g.nv, s, &red, &M);
  if (rc) return rc;
  return done(c, mqpiv::pp_lmm(d->dtype, g, VF, M, ratio, out, removed, s), "local magnitude mask");
}
