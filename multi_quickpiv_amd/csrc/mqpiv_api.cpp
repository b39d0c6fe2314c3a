// mqpiv_api.cpp -- host runtime behind the C ABI declared in include/mqpiv.h.
//
// A context owns a HIP stream and grow-only device staging buffers for the
// host-buffer entry point.  Parameter validation mirrors what the reference
// accepts (src/piv_parameters.jl:29-88, src/multi_quickPIV.jl:40-50) and turns
// its exceptions into status codes with a message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mqpiv.h"
#include "mqpiv_internal.h"

namespace mqpiv {
int fft_len_2d(int fp64, int ss);
size_t lds_2d(int fp64, int L, int alg, int ns0, int ns1, int ss0);
hipError_t launch_piv2d(int fp64, int L, int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix,
                        int n_jobs, int64_t n_ia_total, int ns0, int ns1, int ss0, hipStream_t s);
size_t sat_bytes(const int64_t* dims, int nd);
size_t sat_tmp_bytes(const int64_t* dims, int nd);
hipError_t build_sat(const void* img, int dt, int power, const int64_t* dims, int nd, double* S, double* tmp,
                     hipStream_t s);
int fft_len_3d(int fp64, int ss);
size_t scratch_3d_bytes(int fp64, int L, const Job& j);
hipError_t launch_piv3d(int fp64, int L, const Job& j, void* scratch, const double* sat1_img1, hipStream_t s);
}  // namespace mqpiv

using namespace mqpiv;

struct mqpiv_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  std::string err;
  std::mutex mu;
  // staging for mqpiv_piv (host buffers)
  void* d_stage = nullptr;
  size_t stage_bytes = 0;
  // scratch for multi-kernel (3-D) paths
  void* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  // FP64 integral arrays of the search image (sum, sum of squares) + scan scratch
  void* d_sat = nullptr;
  size_t sat_cap = 0;
  // measurement hooks
  bool profiling = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};

// Brackets one correlation-kernel launch with events when profiling is on.
struct KernelTimer {
  mqpiv_ctx* c;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  KernelTimer(mqpiv_ctx* c_, hipStream_t s_) : c(c_), s(s_) {
    if (c->profiling && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      hipEventRecord(a, s);
  }
  ~KernelTimer() {
    if (a && b) {
      hipEventRecord(b, s);
      c->events.emplace_back(a, b);
    }
  }
};

static int fail(mqpiv_ctx* c, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));

// Grow-only device buffer owned by the context.
static int ensure(mqpiv_ctx* c, void** buf, size_t* cap, size_t need, hipStream_t s) {
  if (need <= *cap) return MQPIV_OK;
  if (*buf) {
    hipStreamSynchronize(s);
    hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
  }
  hipError_t e = hipMalloc(buf, need);
  if (e != hipSuccess)
    return fail(c, MQPIV_ERR_OOM, "hipMalloc(%zu) failed: %s", need, hipGetErrorString(e));
  *cap = need;
  return MQPIV_OK;
}

static int fail(mqpiv_ctx* c, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
static int fail(mqpiv_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIP_OR_FAIL(ctx, expr)                                                              \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(ctx, e_ == hipErrorOutOfMemory ? MQPIV_ERR_OOM : MQPIV_ERR_DEVICE,        \
                  "%s failed: %s", #expr, hipGetErrorString(e_));                           \
  } while (0)

extern "C" const char* mqpiv_version(void) { return "mqpiv 0.1.0 (gfx950)"; }

extern "C" int mqpiv_ctx_create(int device, uint64_t hbm_budget_bytes, mqpiv_ctx** out) {
  if (!out) return MQPIV_ERR_NULL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MQPIV_ERR_DEVICE;
  if (device < 0 || device >= n) return MQPIV_ERR_BAD_PARAMS;
  if (hipSetDevice(device) != hipSuccess) return MQPIV_ERR_DEVICE;
  mqpiv_ctx* c = new mqpiv_ctx();
  c->device = device;
  c->budget = hbm_budget_bytes;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MQPIV_ERR_DEVICE;
  }
  *out = c;
  return MQPIV_OK;
}

extern "C" int mqpiv_ctx_destroy(mqpiv_ctx* c) {
  if (!c) return MQPIV_ERR_NULL;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->d_stage) hipFree(c->d_stage);
  if (c->d_scratch) hipFree(c->d_scratch);
  if (c->d_sat) hipFree(c->d_sat);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return MQPIV_OK;
}

extern "C" const char* mqpiv_last_error(const mqpiv_ctx* c) { return c ? c->err.c_str() : "null context"; }

static int64_t vf_len(int64_t dim, int64_t is, int64_t st) {
  int64_t stop = dim - is + 1;
  return stop >= 1 ? (stop - 1) / st + 1 : 0;
}

static int check_params(mqpiv_ctx* c, const mqpiv_params* p, const int64_t* dims) {
  if (!p || !dims) return fail(c, MQPIV_ERR_NULL, "params/dims is NULL");
  if (p->ndims != 2 && p->ndims != 3) return fail(c, MQPIV_ERR_BAD_PARAMS, "ndims must be 2 or 3 (got %d)", p->ndims);
  if (p->corr_alg < MQPIV_FFTCC || p->corr_alg > MQPIV_MASK_NSQECC)
    return fail(c, MQPIV_ERR_BAD_PARAMS, "unknown corr_alg %d", p->corr_alg);
  if (p->precision_mode != MQPIV_PLAIN_FP32 && p->precision_mode != MQPIV_FP64 &&
      p->precision_mode != MQPIV_AUTO)
    return fail(c, MQPIV_ERR_UNSUPPORTED, "precision_mode %d is not built", p->precision_mode);
  for (int d = 0; d < p->ndims; ++d) {
    if (dims[d] < 1) return fail(c, MQPIV_ERR_BAD_PARAMS, "dims[%d] = %lld", d, (long long)dims[d]);
    if (p->inter_size[d] < 1) return fail(c, MQPIV_ERR_BAD_PARAMS, "interSize[%d] < 1", d);
    if (p->search_margin[d] < 0) return fail(c, MQPIV_ERR_BAD_PARAMS, "searchMargin[%d] < 0", d);
    if (p->step[d] < 1) return fail(c, MQPIV_ERR_BAD_PARAMS, "step[%d] < 1", d);
    if (p->inter_size[d] > (1 << 20) || p->search_margin[d] > (1 << 20))
      return fail(c, MQPIV_ERR_BAD_PARAMS, "window too large");
  }
  return MQPIV_OK;
}

extern "C" int mqpiv_vf_size(const mqpiv_params* p, const int64_t* dims, int64_t* vf, int64_t* n_ia) {
  int rc = check_params(nullptr, p, dims);
  if (rc) return rc;
  int64_t n = 1;
  for (int d = 0; d < 3; ++d) {
    int64_t v = d < p->ndims ? vf_len(dims[d], p->inter_size[d], p->step[d]) : 1;
    if (vf) vf[d] = v;
    n *= v;
  }
  if (n_ia) *n_ia = n;
  return MQPIV_OK;
}

static int build_job(mqpiv_ctx* c, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                     const void* img1, const void* img2, const void* mask, const uint8_t* skip,
                     int32_t out_dtype, void* VF, void* SN, Job* j) {
  int rc = check_params(c, p, dims);
  if (rc) return rc;
  if (in_dtype != MQPIV_F32 && in_dtype != MQPIV_F64) return fail(c, MQPIV_ERR_BAD_PARAMS, "in_dtype %d", in_dtype);
  if (out_dtype != MQPIV_F32 && out_dtype != MQPIV_F64) return fail(c, MQPIV_ERR_BAD_PARAMS, "out_dtype %d", out_dtype);
  if (!img1 || !img2) return fail(c, MQPIV_ERR_NULL, "image pointer is NULL");
  if (p->corr_alg == MQPIV_MASK_NSQECC && !mask) return fail(c, MQPIV_ERR_NULL, "masked NSQECC needs a mask");
  if (p->compute_sn && !SN) return fail(c, MQPIV_ERR_NULL, "computeSN is set but SN is NULL");
  memset(j, 0, sizeof(*j));
  j->img1 = img1;
  j->img2 = img2;
  j->mask = mask;
  j->skip = skip;
  j->vf = VF;
  j->sn = p->compute_sn ? SN : nullptr;
  j->ndims = p->ndims;
  j->n_ia = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < p->ndims) {
      j->dims[d] = dims[d];
      j->is[d] = (int32_t)p->inter_size[d];
      j->sm[d] = (int32_t)p->search_margin[d];
      j->st[d] = (int32_t)std::min<int64_t>(p->step[d], INT32_MAX);
      j->vf_n[d] = (int32_t)vf_len(dims[d], p->inter_size[d], p->step[d]);
    } else {
      j->dims[d] = 1;
      j->is[d] = 1;
      j->sm[d] = 0;
      j->st[d] = 1;
      j->vf_n[d] = 1;
    }
    j->n_ia *= j->vf_n[d];
  }
  j->alg = p->corr_alg;
  j->in_dtype = in_dtype;
  j->out_dtype = out_dtype;
  j->compute_sn = p->compute_sn ? 1 : 0;
  // piv_ops.jl:140-146 (unmasked: threshold >= 0) and :148-157 (masked: > 0)
  if (p->corr_alg == MQPIV_MASK_NSQECC) {
    j->has_threshold = p->threshold > 0;
    j->has_mask_threshold = p->mask_threshold > 0;
  } else {
    j->has_threshold = p->threshold >= 0;
    j->has_mask_threshold = 0;
  }
  j->threshold = p->threshold;
  j->mask_threshold = p->mask_threshold;
  j->precision_mode = p->precision_mode;
  if (!VF && j->n_ia > 0) return fail(c, MQPIV_ERR_NULL, "VF is NULL");
  return MQPIV_OK;
}

// Integral arrays the normalisation reads: NSQECC needs sum(G^2), ZNCC also sum(G).
// (3-D ZNCC also reads the means from integral arrays: sum(F) of image 1.)
static int prepare_sats(mqpiv_ctx* c, Job& j, hipStream_t s) {
  j.sat1 = j.sat2 = j.sat1f = nullptr;
  if (j.alg != ALG_ZNCC && j.alg != ALG_NSQECC) return MQPIV_OK;
  const size_t one = (sat_bytes(j.dims, j.ndims) + 255) & ~(size_t)255;
  const size_t tmp = (sat_tmp_bytes(j.dims, j.ndims) + 255) & ~(size_t)255;
  const int nsat = j.alg == ALG_ZNCC ? (j.ndims == 3 ? 3 : 2) : 1;
  int rc = ensure(c, &c->d_sat, &c->sat_cap, nsat * one + tmp, s);
  if (rc) return rc;
  char* b = (char*)c->d_sat;
  double* t = (double*)(b + nsat * one);
  hipError_t e = build_sat(j.img2, j.in_dtype, 2, j.dims, j.ndims, (double*)b, t, s);
  if (e == hipSuccess && nsat >= 2) e = build_sat(j.img2, j.in_dtype, 1, j.dims, j.ndims, (double*)(b + one), t, s);
  if (e == hipSuccess && nsat >= 3)
    e = build_sat(j.img1, j.in_dtype, 1, j.dims, j.ndims, (double*)(b + 2 * one), t, s);
  if (e != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "integral array build failed: %s", hipGetErrorString(e));
  j.sat2 = (const double*)b;
  j.sat1 = nsat >= 2 ? (const double*)(b + one) : nullptr;
  j.sat1f = nsat >= 3 ? (const double*)(b + 2 * one) : nullptr;
  return MQPIV_OK;
}

static int launch_job(mqpiv_ctx* c, Job& j, hipStream_t s) {
  if (j.n_ia == 0) return MQPIV_OK;
  int src = prepare_sats(c, j, s);
  if (src) return src;
  if (j.ndims == 2) {
    const int ss0 = j.is[0] + 2 * j.sm[0], ss1 = j.is[1] + 2 * j.sm[1];
    const int ss = std::max(ss0, ss1);
    const int ns0 = 2 * j.sm[0] + 1, ns1 = 2 * j.sm[1] + 1;
    auto fits = [&](int fp64) {
      int L = fft_len_2d(fp64, ss);
      return L > 0 && lds_2d(fp64, L, j.alg, ns0, ns1, ss0) <= 160 * 1024;
    };
    int fp64 = 0;
    if (j.precision_mode == MQPIV_FP64) {
      if (!fits(1))
        return fail(c, MQPIV_ERR_UNSUPPORTED, "FP64 transforms support 2-D search sizes up to 96 (got %d)", ss);
      fp64 = 1;
    } else if (j.precision_mode == MQPIV_AUTO) {
      bool want = j.alg == ALG_MASK_NSQECC || (j.alg == ALG_ZNCC && j.compute_sn);
      fp64 = want && fits(1);
    }
    if (!fits(fp64))
      return fail(c, MQPIV_ERR_UNSUPPORTED, "search size %d exceeds the largest 2-D transform (128)", ss);
    const int L = fft_len_2d(fp64, ss);
    hipError_t e;
    {
      KernelTimer kt(c, s);
      e = launch_piv2d(fp64, L, j.alg, j, nullptr, nullptr, 1, j.n_ia, ns0, ns1, ss0, s);
    }
    if (e != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "2-D launch failed: %s", hipGetErrorString(e));
    return MQPIV_OK;
  }
  // 3-D
  if (j.alg == ALG_MASK_NSQECC) return fail(c, MQPIV_ERR_UNSUPPORTED, "masked NSQECC is built for 2-D only");
  const int ss = std::max({j.is[0] + 2 * j.sm[0], j.is[1] + 2 * j.sm[1], j.is[2] + 2 * j.sm[2]});
  int fp64 = 0;
  if (j.precision_mode == MQPIV_FP64) fp64 = 1;
  else if (j.precision_mode == MQPIV_AUTO) fp64 = (j.alg == ALG_ZNCC && j.compute_sn) ? 1 : 0;
  const int L = fft_len_3d(fp64, ss);
  if (L < 0) return fail(c, MQPIV_ERR_UNSUPPORTED, "3-D search size %d exceeds the largest transform (64)", ss);
  int rc = ensure(c, &c->d_scratch, &c->scratch_bytes, scratch_3d_bytes(fp64, L, j), s);
  if (rc) return rc;
  hipError_t e;
  {
    KernelTimer kt(c, s);
    e = launch_piv3d(fp64, L, j, c->d_scratch, j.sat1f, s);
  }
  if (e != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "3-D launch failed: %s", hipGetErrorString(e));
  return MQPIV_OK;
}

extern "C" int mqpiv_piv_device(mqpiv_ctx* c, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                                const void* d_img1, const void* d_img2, const void* d_mask,
                                const uint8_t* d_skip, int32_t out_dtype, void* d_VF, void* d_SN,
                                void* stream) {
  if (!c) return MQPIV_ERR_NULL;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "hipSetDevice failed");
  Job j;
  int rc = build_job(c, p, dims, in_dtype, d_img1, d_img2, d_mask, d_skip, out_dtype, d_VF, d_SN, &j);
  if (rc) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  return launch_job(c, j, s);
}

extern "C" int mqpiv_piv(mqpiv_ctx* c, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                         const void* img1, const void* img2, const void* mask, const uint8_t* skip,
                         int32_t out_dtype, void* VF, void* SN) {
  return mqpiv_piv_ex(c, p, dims, in_dtype, img1, img2, mask, skip, out_dtype, VF, SN, nullptr);
}

extern "C" int mqpiv_piv_ex(mqpiv_ctx* c, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                            const void* img1, const void* img2, const void* mask, const uint8_t* skip,
                            int32_t out_dtype, void* VF, void* SN, int32_t* peaks) {
  if (!c) return MQPIV_ERR_NULL;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "hipSetDevice failed");
  Job j;
  int rc = build_job(c, p, dims, in_dtype, img1, img2, mask, skip, out_dtype, VF, SN, &j);
  if (rc) return rc;
  const int64_t npx = j.dims[0] * j.dims[1] * j.dims[2];
  const size_t esz = in_dtype == MQPIV_F32 ? 4 : 8;
  const size_t osz = out_dtype == MQPIV_F32 ? 4 : 8;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t b_img = al(npx * esz);
  const size_t b_mask = (p->corr_alg == MQPIV_MASK_NSQECC) ? b_img : 0;
  const size_t b_skip = skip ? al(j.n_ia) : 0;
  const size_t b_vf = al(j.n_ia * j.ndims * osz);
  const size_t b_sn = j.sn ? al(j.n_ia * osz) : 0;
  const size_t b_pk = peaks ? al(j.n_ia * j.ndims * sizeof(int32_t)) : 0;
  const size_t total = 2 * b_img + b_mask + b_skip + b_vf + b_sn + b_pk;
  if (c->budget && total > c->budget)
    return fail(c, MQPIV_ERR_OOM, "request needs %zu bytes, over the HBM budget %llu", total,
                (unsigned long long)c->budget);
  if (total > c->stage_bytes) {
    if (c->d_stage) {
      hipStreamSynchronize(c->stream);
      hipFree(c->d_stage);
      c->d_stage = nullptr;
      c->stage_bytes = 0;
    }
    HIP_OR_FAIL(c, hipMalloc(&c->d_stage, total));
    c->stage_bytes = total;
  }
  char* b = (char*)c->d_stage;
  char* d1 = b;
  char* d2 = d1 + b_img;
  char* dm = d2 + b_img;
  char* dk = dm + b_mask;
  char* dv = dk + b_skip;
  char* ds = dv + b_vf;
  char* dp = ds + b_sn;
  hipStream_t s = c->stream;
  HIP_OR_FAIL(c, hipMemcpyAsync(d1, img1, npx * esz, hipMemcpyHostToDevice, s));
  HIP_OR_FAIL(c, hipMemcpyAsync(d2, img2, npx * esz, hipMemcpyHostToDevice, s));
  if (b_mask) HIP_OR_FAIL(c, hipMemcpyAsync(dm, mask, npx * esz, hipMemcpyHostToDevice, s));
  if (b_skip) HIP_OR_FAIL(c, hipMemcpyAsync(dk, skip, j.n_ia, hipMemcpyHostToDevice, s));
  j.img1 = d1;
  j.img2 = d2;
  j.mask = b_mask ? dm : nullptr;
  j.skip = b_skip ? (const uint8_t*)dk : nullptr;
  j.vf = dv;
  j.sn = b_sn ? ds : nullptr;
  j.peaks = b_pk ? (int32_t*)dp : nullptr;
  rc = launch_job(c, j, s);
  if (rc) return rc;
  if (j.n_ia > 0) {
    HIP_OR_FAIL(c, hipMemcpyAsync(VF, dv, j.n_ia * j.ndims * osz, hipMemcpyDeviceToHost, s));
    if (b_sn) HIP_OR_FAIL(c, hipMemcpyAsync(SN, ds, j.n_ia * osz, hipMemcpyDeviceToHost, s));
    if (b_pk)
      HIP_OR_FAIL(c, hipMemcpyAsync(peaks, dp, j.n_ia * j.ndims * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  }
  HIP_OR_FAIL(c, hipStreamSynchronize(s));
  return MQPIV_OK;
}

extern "C" int mqpiv_profile_begin(mqpiv_ctx* c) {
  if (!c) return MQPIV_ERR_NULL;
  std::lock_guard<std::mutex> lk(c->mu);
  for (auto& e : c->events) {
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  c->events.clear();
  c->profiling = true;
  return MQPIV_OK;
}

extern "C" int mqpiv_profile_end(mqpiv_ctx* c, double* kernel_ms, int32_t cap, int32_t* n) {
  if (!c) return MQPIV_ERR_NULL;
  std::lock_guard<std::mutex> lk(c->mu);
  c->profiling = false;
  int32_t k = 0;
  for (auto& e : c->events) {
    hipEventSynchronize(e.second);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e.first, e.second);
    if (kernel_ms && k < cap) kernel_ms[k] = ms;
    ++k;
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  c->events.clear();
  if (n) *n = k;
  return MQPIV_OK;
}
