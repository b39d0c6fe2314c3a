// mqpiv_api.cpp -- host runtime behind the C ABI declared in include/mqpiv.h.
//
// A context owns a HIP stream and grow-only device staging buffers for the
// host-buffer entry point.  Parameter validation mirrors what the reference
// accepts (src/piv_parameters.jl:29-88, src/multi_quickPIV.jl:40-50) and turns
// its exceptions into status codes with a message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>
#include <vector>

#include "../../include/mqpiv.h"
#include "mqpiv_batch.h"
#include "mqpiv_ctx.h"
#include "mqpiv_internal.h"

static_assert(MQPIV_SPLIT_MFMA == mqpiv::kPrecSplitMfma, "precision mode mirror");

namespace mqpiv {
int fft_len_2d(int fp64, int ss);
size_t lds_2d(int fp64, int L, int alg, int ns0, int ns1, int ss0);
hipError_t launch_piv2d(int fp64, int L, int alg, const Job& j0, const Job* d_jobs, const int64_t* d_prefix,
                        int n_jobs, int64_t n_ia_total, int ns0, int ns1, int ss0, void* gz, hipStream_t s, int Lr);
size_t gz_scratch_2d(int fp64, int L, int64_t n_ia);
int wave2d_len(int alg, int is0, int is1, int ss, int ns0, int ns1);
size_t sat_bytes(const int64_t* dims, int nd);
size_t sat_tmp_bytes(const int64_t* dims, int nd);
hipError_t build_sat(const void* img, int dt, int power, const int64_t* dims, int nd, double* S, double* tmp,
                     hipStream_t s);
hipError_t build_sat_pair(const void* img, int dt, const int64_t* dims, int nd, double* S2, double* S1,
                          double* tmp, hipStream_t s);
int fft_len_3d(int fp64, int ss);
size_t scratch_3d_bytes(int fp64, int L, const Job& j);
hipError_t launch_piv3d(int fp64, int L, const Job& j, void* scratch, const double* sat1_img1, hipStream_t s);
hipError_t launch_skip_flags(const Job& j, const void* mask_vf, uint8_t* flags, hipStream_t s);
size_t means3d_bytes(const Job& j);
hipError_t launch_means3d(const Job& j, void* buf, hipStream_t s);
hipError_t launch_coarse_to_shift(const double* vfc, const int32_t* nc, const int32_t* isc, const int32_t* stc,
                                  const int32_t* nf, const int32_t* isf, const int32_t* stf, int nd,
                                  int64_t n_fine, int32_t* out, hipStream_t s);
}  // namespace mqpiv

using namespace mqpiv;

// Grow-only device buffer owned by the context.
int ensure(mqpiv_ctx* c, void** buf, size_t* cap, size_t need, hipStream_t s) {
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

int fail(mqpiv_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

// ---------------------------------------------------------------------------
// Static scratch of the allocation-free entry point (mqpiv_piv_device_noalloc):
// the integral arrays of one 2-D call, in a file-scope device array that each
// GPU holds from library load to process exit.  Host code never uses the
// array's name as an address: its per-device address comes from
// hipGetSymbolAddress at context creation (with that device current).  The
// integral-array build writes every entry, its zero border included, so
// nothing in it has to start at zero.
static __device__ __align__(256) unsigned char scratch[MQPIV_NOALLOC_SCRATCH_BYTES];
static constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::string g_scratch_err[kMaxDevices];
static std::mutex g_scratch_mu;

static void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e == hipSuccess && ptr) g_scratch_ptr[device] = ptr;
  else g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
}

// Smallest sizes the context's grow-only refine/transform buffer and job tables
// take (64 KiB of tables: about 300 jobs).
static constexpr size_t kMinGrowBytes = (size_t)1 << 20;
static constexpr size_t kMinTableBytes = (size_t)64 << 10;

// The context's global transform buffers (gz_scratch_2d bytes): a new
// allocation gets its slot flags cleared (launches leave them clear).
static int ensure_gz(mqpiv_ctx* c, size_t need, hipStream_t s) {
  const size_t cap0 = c->gz_cap;
  // at least 1 MiB (a refine list of 262,080 IAs), so that a context that
  // serves requests of growing size rarely reallocates (each regrow waits for
  // the stream)
  int rc = ensure(c, &c->d_gz, &c->gz_cap, std::max(need, kMinGrowBytes), s);
  if (rc) return rc;
  if (c->gz_cap != cap0) HIP_OR_FAIL(c, hipMemsetAsync(c->d_gz, 0, kGzFlagBytes, s));
  return MQPIV_OK;
}

extern "C" const char* mqpiv_version(void) { return "mqpiv 0.1.0 (gfx950)"; }

extern "C" int mqpiv_ctx_create(int device, uint64_t hbm_budget_bytes, mqpiv_ctx** out) {
  if (!out) return MQPIV_ERR_NULL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MQPIV_ERR_DEVICE;
  if (device < 0 || device >= n) return MQPIV_ERR_BAD_PARAMS;
  if (hipSetDevice(device) != hipSuccess) return MQPIV_ERR_DEVICE;
  noalloc_setup(device);  // the allocation-free entry point's scratch address on this device
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
  if (c->d_bsat) hipFree(c->d_bsat);
  if (c->d_table) hipFree(c->d_table);
  if (c->h_table) hipHostFree(c->h_table);
  if (c->d_pp) hipFree(c->d_pp);
  if (c->d_mp) hipFree(c->d_mp);
  if (c->d_gz) hipFree(c->d_gz);
  if (c->d_flags) hipFree(c->d_flags);
  if (c->table_done) hipEventDestroy(c->table_done);
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
  if (p->precision_mode != MQPIV_PLAIN_FP32 && p->precision_mode != MQPIV_SPLIT_MFMA &&
      p->precision_mode != MQPIV_FP64 && p->precision_mode != MQPIV_AUTO)
    return fail(c, MQPIV_ERR_UNSUPPORTED, "precision_mode %d is not built", p->precision_mode);
  if (p->precision_mode == MQPIV_SPLIT_MFMA && (p->ndims != 2 || p->corr_alg == MQPIV_MASK_NSQECC))
    return fail(c, MQPIV_ERR_UNSUPPORTED, "the split-precision matrix-core form runs 2-D FFTCC, ZNCC and NSQECC");
  for (int d = 0; d < p->ndims; ++d) {
    if (dims[d] < 1) return fail(c, MQPIV_ERR_BAD_PARAMS, "dims[%d] = %lld", d, (long long)dims[d]);
    if (p->inter_size[d] < 1) return fail(c, MQPIV_ERR_BAD_PARAMS, "interSize[%d] < 1", d);
    if (p->search_margin[d] < 0) return fail(c, MQPIV_ERR_BAD_PARAMS, "searchMargin[%d] < 0", d);
    if (p->step[d] < 1) return fail(c, MQPIV_ERR_BAD_PARAMS, "step[%d] < 1", d);
    if (p->inter_size[d] > (1 << 20) || p->search_margin[d] > (1 << 20))
      return fail(c, MQPIV_ERR_BAD_PARAMS, "window too large");
  }
  // the kernels index IAs and image coordinates with 32-bit integers
  int64_t n_ia = 1;
  for (int d = 0; d < p->ndims; ++d) {
    if (dims[d] >= ((int64_t)1 << 31))
      return fail(c, MQPIV_ERR_UNSUPPORTED, "dims[%d] = %lld is 2^31 or more", d, (long long)dims[d]);
    n_ia *= vf_len(dims[d], p->inter_size[d], p->step[d]);
    if (n_ia >= ((int64_t)1 << 31))
      return fail(c, MQPIV_ERR_UNSUPPORTED, "2^31 or more IAs in one call; split the frame");
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

// Integral arrays the normalisation reads: 2-D NSQECC needs sum(G^2), 2-D ZNCC
// also sum(G); 2-D ZNCC with FP64 transforms needs none (it forms its window
// sums in the kernel).  3-D forms its window sums from the slabs (K1), so 3-D
// ZNCC needs only the means' sums of image 2 and image 1, and 3-D NSQECC none.
// Integral arrays a job needs: count, bytes of one (256-aligned) and of the
// prefix-scan temporary.
static int sat_count(int alg, int ndims, int fp64 = 0) {
  if (ndims == 3) return alg == ALG_ZNCC ? 2 : 0;
  if (alg == ALG_ZNCC) return fp64 == 1 ? 0 : 2;
  return alg == ALG_NSQECC ? 1 : 0;
}

// 3-D ZNCC with unshifted search regions takes its means from per-IA sums
// (launch_means3d) instead of the two integral arrays.
static bool means3d_job(const Job& j) { return j.ndims == 3 && j.alg == ALG_ZNCC && !j.shift; }

// Workspace a job's integral arrays (or 3-D mean sums) take (as request_ws_bytes).
static size_t job_ws_bytes(const Job& j, int fp64) {
  if (means3d_job(j)) return means3d_bytes(j);
  const int nsat = sat_count(j.alg, j.ndims, fp64);
  if (!nsat) return 0;
  const size_t one = (sat_bytes(j.dims, j.ndims) + 255) & ~(size_t)255;
  const size_t tmp = (sat_tmp_bytes(j.dims, j.ndims) + 255) & ~(size_t)255;
  return nsat * one + tmp;
}

size_t request_ws_bytes(const mqpiv_request& r) {
  const int nd = r.params.ndims;
  if (nd == 3 && r.params.corr_alg == MQPIV_ZNCC) {  // requests carry no shifts
    Job j{};
    j.ndims = 3;
    j.alg = ALG_ZNCC;
    j.n_ia = 1;
    for (int d = 0; d < 3; ++d) {
      j.dims[d] = r.dims[d];
      j.vf_n[d] = (int32_t)vf_len(r.dims[d], r.params.inter_size[d], r.params.step[d]);
      j.n_ia *= j.vf_n[d];
    }
    return means3d_bytes(j);
  }
  const int nsat = sat_count(r.params.corr_alg, nd);
  if (!nsat) return 0;
  const size_t one = (sat_bytes(r.dims, nd) + 255) & ~(size_t)255;
  const size_t tmp = (sat_tmp_bytes(r.dims, nd) + 255) & ~(size_t)255;
  return nsat * one + tmp;
}

// Builds the job's integral arrays in ws (request_ws_bytes of it) or, when ws is
// NULL, in the context's grow-only buffer.
static int prepare_sats(mqpiv_ctx* c, Job& j, hipStream_t s, char* ws = nullptr, int fp64 = 0) {
  j.sat1 = j.sat2 = j.sat1f = nullptr;
  const int nsat = sat_count(j.alg, j.ndims, fp64);
  if (!nsat) return MQPIV_OK;
  if (means3d_job(j)) {  // the per-IA sums of the means, by separable passes
    if (!ws) {
      int rc = ensure(c, &c->d_sat, &c->sat_cap, means3d_bytes(j), s);
      if (rc) return rc;
    }
    char* b = ws ? ws : (char*)c->d_sat;
    hipError_t e = launch_means3d(j, b, s);
    if (e != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "3-D mean sums failed: %s", hipGetErrorString(e));
    j.sat1f = (const double*)b;
    return MQPIV_OK;
  }
  const size_t one = (sat_bytes(j.dims, j.ndims) + 255) & ~(size_t)255;
  const size_t tmp = (sat_tmp_bytes(j.dims, j.ndims) + 255) & ~(size_t)255;
  if (!ws) {
    int rc = ensure(c, &c->d_sat, &c->sat_cap, nsat * one + tmp, s);
    if (rc) return rc;
  }
  char* b = ws ? ws : (char*)c->d_sat;
  double* t = (double*)(b + nsat * one);
  if (j.ndims == 3) {  // ZNCC with shifted regions: sums of image 2 and of image 1 (the means)
    hipError_t e = build_sat(j.img2, j.in_dtype, 1, j.dims, 3, (double*)b, t, s);
    if (e == hipSuccess) e = build_sat(j.img1, j.in_dtype, 1, j.dims, 3, (double*)(b + one), t, s);
    if (e != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "integral array build failed: %s", hipGetErrorString(e));
    j.sat1 = (const double*)b;
    j.sat1f = (const double*)(b + one);
    return MQPIV_OK;
  }
  hipError_t e = nsat >= 2 ? build_sat_pair(j.img2, j.in_dtype, j.dims, j.ndims, (double*)b, (double*)(b + one), t, s)
                           : build_sat(j.img2, j.in_dtype, 2, j.dims, j.ndims, (double*)b, t, s);
  if (e != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "integral array build failed: %s", hipGetErrorString(e));
  j.sat2 = (const double*)b;
  j.sat1 = nsat >= 2 ? (const double*)(b + one) : nullptr;
  return MQPIV_OK;
}

// AUTO's 3-D precision: FP64 transforms where FP32 ones miss the parity bar
// (ZNCC with SN; masked NSQECC, whose num = sum(F^2) + C2 - 2 C1 cancels near the
// match), FP32 otherwise.
static int auto_fp64_3d(const Job& j) {
  return ((j.alg == ALG_ZNCC && j.compute_sn) || j.alg == ALG_MASK_NSQECC) ? 1 : 0;
}

// FP64 transform length that recomputes the flagged IAs of a precision-3 job.
static int refine_len(const Job& j) {
  return fft_len_2d(1, std::max(j.is[0] + 2 * j.sm[0], j.is[1] + 2 * j.sm[1]));
}

// Precision and transform length of a 2-D job (see MQPIV_AUTO in mqpiv.h).
static int choose_2d(mqpiv_ctx* c, const Job& j, int* fp64_out, int* L_out) {
  const int ss0 = j.is[0] + 2 * j.sm[0], ss1 = j.is[1] + 2 * j.sm[1];
  const int ss = std::max(ss0, ss1);
  const int ns0 = 2 * j.sm[0] + 1, ns1 = 2 * j.sm[1] + 1;
  // the 2-D kernels address a window (image, or FP64 integral array with rows
  // of dims[0] + 1) with 32-bit byte offsets: at most 128 rows of (dims[0] + 1)
  // entries of at most 8 bytes, plus one row
  if (((int64_t)j.dims[0] + 2) * 1024 >= ((int64_t)1 << 32))
    return fail(c, MQPIV_ERR_UNSUPPORTED, "dims[0] = %lld is too large for 32-bit in-window offsets",
                (long long)j.dims[0]);
  auto fits = [&](int fp64) {
    int L = fft_len_2d(fp64, ss);
    return L > 0 && lds_2d(fp64, L, j.alg, ns0, ns1, ss0) <= 160 * 1024;
  };
  int fp64 = 0;
  if (j.precision_mode == MQPIV_SPLIT_MFMA) {
    const int L = wave2d_len(j.alg, j.is[0], j.is[1], ss, ns0, ns1);
    if (L < 0)
      return fail(c, MQPIV_ERR_UNSUPPORTED,
                  "the split-precision matrix-core form takes FFTCC, ZNCC and NSQECC with interSize <= 32 and search "
                  "size <= 64, or interSize <= 64 and searchMargin <= 16, windows up to 48 x 48 (got interSize %d x "
                  "%d, searchMargin %d x %d)",
                  j.is[0], j.is[1], j.sm[0], j.sm[1]);
    *fp64_out = 2;  // the matrix-core kernel (launch_piv2d)
    *L_out = L;
    return MQPIV_OK;
  }
  if (j.precision_mode == MQPIV_FP64) {
    if (!fits(1))
      return fail(c, MQPIV_ERR_UNSUPPORTED, "FP64 transforms support 2-D search sizes up to 128 (got %d)", ss);
    fp64 = 1;
  } else if (j.precision_mode == MQPIV_AUTO) {
    bool want = j.alg == ALG_MASK_NSQECC || (j.alg == ALG_ZNCC && j.compute_sn);
    fp64 = want && fits(1);
    // masked NSQECC: FP32 transforms, the IAs they cannot settle recomputed by
    // the FP64 transforms of the same length (precision 4)
    if (fp64 && j.alg == ALG_MASK_NSQECC && !j.shift && fits(0) && fft_len_2d(0, ss) == fft_len_2d(1, ss)) {
      *fp64_out = 4;
      *L_out = fft_len_2d(0, ss);
      return MQPIV_OK;
    }
    // where FP32-level transforms meet the bar, the split-precision matrix-core
    // form when it takes the job (it runs faster than the FP32 transforms); for
    // ZNCC with SN, that form with the IAs it flags recomputed by the FP64
    // transforms (precision 3)
    if (!j.shift && (!fp64 || j.alg == ALG_ZNCC)) {
      const int Lw = wave2d_len(j.alg, j.is[0], j.is[1], ss, ns0, ns1);
      if (Lw > 0) {
        *fp64_out = fp64 ? 3 : 2;
        *L_out = Lw;
        return MQPIV_OK;
      }
    }
  }
  if (!fits(fp64))
    return fail(c, MQPIV_ERR_UNSUPPORTED, "search size %d exceeds the largest 2-D transform (128)", ss);
  *fp64_out = fp64;
  *L_out = fft_len_2d(fp64, ss);
  return MQPIV_OK;
}

extern "C" int mqpiv_transform_form(const mqpiv_params* p, const int64_t* dims, int32_t* form, int32_t* length) {
  if (!form || !length) return MQPIV_ERR_NULL;
  int rc = check_params(nullptr, p, dims);
  if (rc) return rc;
  Job j;
  memset(&j, 0, sizeof(j));
  j.ndims = p->ndims;
  for (int d = 0; d < 3; ++d) {
    j.dims[d] = d < p->ndims ? dims[d] : 1;
    j.is[d] = d < p->ndims ? (int32_t)p->inter_size[d] : 1;
    j.sm[d] = d < p->ndims ? (int32_t)p->search_margin[d] : 0;
  }
  j.alg = p->corr_alg;
  j.compute_sn = p->compute_sn ? 1 : 0;
  j.precision_mode = p->precision_mode;
  if (p->ndims == 3) {
    const int ss = std::max({j.is[0] + 2 * j.sm[0], j.is[1] + 2 * j.sm[1], j.is[2] + 2 * j.sm[2]});
    int fp64 = 0;
    if (j.precision_mode == MQPIV_FP64) fp64 = 1;
    else if (j.precision_mode == MQPIV_AUTO) fp64 = auto_fp64_3d(j);
    const int L = fft_len_3d(fp64, ss);
    if (L < 0) return MQPIV_ERR_UNSUPPORTED;
    *form = fp64;
    *length = L;
    return MQPIV_OK;
  }
  int fp64 = 0, L = 0;
  rc = choose_2d(nullptr, j, &fp64, &L);
  if (rc) return rc;
  *form = fp64;
  *length = L;
  return MQPIV_OK;
}

// ws: workspace for the integral arrays (else the context's buffer).  form /
// gz (2-D, the allocation-free entry point): the precision and length to run
// and the buffer of gz_scratch_2d bytes for them, instead of choose_2d's and the
// context's.
static int launch_job(mqpiv_ctx* c, Job& j, hipStream_t s, char* ws = nullptr, const int* form = nullptr,
                      char* gz = nullptr) {
  if (j.n_ia == 0) return MQPIV_OK;
  if (j.ndims == 2) {
    const int ss0 = j.is[0] + 2 * j.sm[0];
    const int ns0 = 2 * j.sm[0] + 1, ns1 = 2 * j.sm[1] + 1;
    int fp64 = 0, L = 0;
    int rc2 = MQPIV_OK;
    if (form) {
      fp64 = form[0];
      L = form[1];
    } else {
      rc2 = choose_2d(c, j, &fp64, &L);
    }
    if (rc2) return rc2;
    rc2 = prepare_sats(c, j, s, ws, fp64);
    if (rc2) return rc2;
    const size_t gzb = gz_scratch_2d(fp64, L, j.n_ia);
    if (gzb && !gz) {
      rc2 = ensure_gz(c, gzb, s);
      if (rc2) return rc2;
      gz = (char*)c->d_gz;
    }
    hipError_t e;
    {
      KernelTimer kt(c, s);
      e = launch_piv2d(fp64, L, j.alg, j, nullptr, nullptr, 1, j.n_ia, ns0, ns1, ss0, gz, s,
                       fp64 >= 3 ? refine_len(j) : 0);
    }
    if (e != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "2-D launch failed: %s", hipGetErrorString(e));
    return MQPIV_OK;
  }
  // 3-D
  int src = prepare_sats(c, j, s, ws);
  if (src) return src;
  const int ss = std::max({j.is[0] + 2 * j.sm[0], j.is[1] + 2 * j.sm[1], j.is[2] + 2 * j.sm[2]});
  int fp64 = 0;
  if (j.precision_mode == MQPIV_FP64) fp64 = 1;
  else if (j.precision_mode == MQPIV_AUTO) fp64 = auto_fp64_3d(j);
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

extern "C" int mqpiv_piv_device_noalloc(mqpiv_ctx* c, const mqpiv_params* p, const int64_t* dims,
                                        int32_t in_dtype, const void* d_img1, const void* d_img2,
                                        const void* d_mask, const uint8_t* d_skip, int32_t out_dtype, void* d_VF,
                                        void* d_SN, void* stream) {
  if (!c) return MQPIV_ERR_NULL;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "hipSetDevice failed");
  Job j;
  int rc = build_job(c, p, dims, in_dtype, d_img1, d_img2, d_mask, d_skip, out_dtype, d_VF, d_SN, &j);
  if (rc) return rc;
  if (j.ndims != 2)
    return fail(c, MQPIV_ERR_UNSUPPORTED, "the allocation-free entry point runs 2-D pairs (3-D needs transform scratch)");
  int fp64 = 0, L = 0;
  rc = choose_2d(c, j, &fp64, &L);
  if (rc) return rc;
  // precision 3 keeps its refine list after the integral arrays; where the two
  // do not fit, the FP64 transforms run every IA (they need no integral arrays)
  if (fp64 == 3 && job_ws_bytes(j, 3) + gz_scratch_2d(3, L, j.n_ia) > MQPIV_NOALLOC_SCRATCH_BYTES) {
    fp64 = 1;
    L = refine_len(j);
  }
  if (fp64 == 4 && L == 128)  // the FP64 recomputation needs global transform buffers there
    return fail(c, MQPIV_ERR_UNSUPPORTED,
                "the allocation-free entry point does not run FP64 transforms at search size 128 (global transform buffers)");
  if (fp64 == 4 && gz_scratch_2d(4, L, j.n_ia) > MQPIV_NOALLOC_SCRATCH_BYTES) fp64 = 1;
  if (fp64 < 3 && gz_scratch_2d(fp64, L, j.n_ia))
    return fail(c, MQPIV_ERR_UNSUPPORTED,
                "the allocation-free entry point does not run FP64 transforms at search size 128 (global transform buffers)");
  void* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk2(g_scratch_mu);
    if (c->device >= 0 && c->device < kMaxDevices) ptr = g_scratch_ptr[c->device];
    if (!ptr)
      return fail(c, MQPIV_ERR_DEVICE, "no static scratch on device %d: %s", c->device,
                  c->device >= 0 && c->device < kMaxDevices ? g_scratch_err[c->device].c_str() : "device index");
  }
  const size_t need = job_ws_bytes(j, fp64);
  if (need > MQPIV_NOALLOC_SCRATCH_BYTES)
    return fail(c, MQPIV_ERR_OOM, "the pair needs %zu bytes of integral arrays; the static scratch holds %llu", need,
                (unsigned long long)MQPIV_NOALLOC_SCRATCH_BYTES);
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const int form[2] = {fp64, L};
  return launch_job(c, j, s, need ? (char*)ptr : nullptr, form, fp64 >= 3 ? (char*)ptr + need : nullptr);
}

extern "C" int mqpiv_piv(mqpiv_ctx* c, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                         const void* img1, const void* img2, const void* mask, const uint8_t* skip,
                         int32_t out_dtype, void* VF, void* SN) {
  return mqpiv_piv_ex(c, p, dims, in_dtype, img1, img2, mask, skip, out_dtype, VF, SN, nullptr);
}

// Host buffers: validates the request, copies the inputs into the context's
// staging buffer, runs `run` on the job with device pointers, copies the outputs
// back and synchronises.
// mask_vf (double-mask calls): the skip filter's mask, staged beside the
// others; the flags it gives (one byte per IA) become the job's skip mask.
template <typename Run>
static int run_staged(mqpiv_ctx* c, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                      const void* img1, const void* img2, const void* mask, const uint8_t* skip,
                      int32_t out_dtype, void* VF, void* SN, int32_t* peaks, Run&& run,
                      const void* mask_vf = nullptr) {
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
  const size_t b_skip = (skip || mask_vf) ? al(j.n_ia) : 0;
  const size_t b_mvf = mask_vf ? b_img : 0;
  const size_t b_vf = al(j.n_ia * j.ndims * osz);
  const size_t b_sn = j.sn ? al(j.n_ia * osz) : 0;
  const size_t b_pk = peaks ? al(j.n_ia * j.ndims * sizeof(int32_t)) : 0;
  const size_t total = 2 * b_img + b_mask + b_mvf + b_skip + b_vf + b_sn + b_pk;
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
  char* dmv = dm + b_mask;
  char* dk = dmv + b_mvf;
  char* dv = dk + b_skip;
  char* ds = dv + b_vf;
  char* dp = ds + b_sn;
  hipStream_t s = c->stream;
  HIP_OR_FAIL(c, hipMemcpyAsync(d1, img1, npx * esz, hipMemcpyHostToDevice, s));
  HIP_OR_FAIL(c, hipMemcpyAsync(d2, img2, npx * esz, hipMemcpyHostToDevice, s));
  if (b_mask) HIP_OR_FAIL(c, hipMemcpyAsync(dm, mask, npx * esz, hipMemcpyHostToDevice, s));
  if (b_mvf) HIP_OR_FAIL(c, hipMemcpyAsync(dmv, mask_vf, npx * esz, hipMemcpyHostToDevice, s));
  if (skip) HIP_OR_FAIL(c, hipMemcpyAsync(dk, skip, j.n_ia, hipMemcpyHostToDevice, s));
  j.img1 = d1;
  j.img2 = d2;
  j.mask = b_mask ? dm : nullptr;
  j.skip = skip ? (const uint8_t*)dk : nullptr;
  if (mask_vf && (j.has_threshold || j.has_mask_threshold)) {
    HIP_OR_FAIL(c, launch_skip_flags(j, dmv, (uint8_t*)dk, s));
    j.skip = (const uint8_t*)dk;
  }
  j.vf = dv;
  j.sn = b_sn ? ds : nullptr;
  j.peaks = b_pk ? (int32_t*)dp : nullptr;
  rc = run(j, s);
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

extern "C" int mqpiv_piv_ex(mqpiv_ctx* c, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                            const void* img1, const void* img2, const void* mask, const uint8_t* skip,
                            int32_t out_dtype, void* VF, void* SN, int32_t* peaks) {
  return run_staged(c, p, dims, in_dtype, img1, img2, mask, skip, out_dtype, VF, SN, peaks,
                    [&](Job& j, hipStream_t s) { return launch_job(c, j, s); });
}

// Checks of the double-mask form: the masked measure, both masks given.
static int check_masks(mqpiv_ctx* c, const mqpiv_params* p, const void* mask_vf, const void* mask_ia) {
  if (!p) return fail(c, MQPIV_ERR_NULL, "params is NULL");
  if (p->corr_alg != MQPIV_MASK_NSQECC)
    return fail(c, MQPIV_ERR_BAD_PARAMS, "the two-mask form runs the masked NSQECC measure");
  if (!mask_vf || !mask_ia) return fail(c, MQPIV_ERR_NULL, "a mask pointer is NULL");
  return MQPIV_OK;
}

extern "C" int mqpiv_piv_masks(mqpiv_ctx* c, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                               const void* img1, const void* img2, const void* mask_vf, const void* mask_ia,
                               int32_t out_dtype, void* VF, void* SN, int32_t* peaks) {
  if (!c) return MQPIV_ERR_NULL;
  int rc = check_masks(c, p, mask_vf, mask_ia);
  if (rc) return rc;
  return run_staged(
      c, p, dims, in_dtype, img1, img2, mask_ia, nullptr, out_dtype, VF, SN, peaks,
      [&](Job& j, hipStream_t s) { return launch_job(c, j, s); }, mask_vf);
}

extern "C" int mqpiv_piv_masks_device(mqpiv_ctx* c, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                                      const void* d_img1, const void* d_img2, const void* d_mask_vf,
                                      const void* d_mask_ia, int32_t out_dtype, void* d_VF, void* d_SN,
                                      void* stream) {
  if (!c) return MQPIV_ERR_NULL;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "hipSetDevice failed");
  int rc = check_masks(c, p, d_mask_vf, d_mask_ia);
  if (rc) return rc;
  Job j;
  rc = build_job(c, p, dims, in_dtype, d_img1, d_img2, d_mask_ia, nullptr, out_dtype, d_VF, d_SN, &j);
  if (rc) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  if (j.n_ia > 0 && (j.has_threshold || j.has_mask_threshold)) {
    rc = ensure(c, &c->d_flags, &c->flags_cap, (size_t)j.n_ia, s);
    if (rc) return rc;
    HIP_OR_FAIL(c, launch_skip_flags(j, d_mask_vf, (uint8_t*)c->d_flags, s));
    j.skip = (const uint8_t*)c->d_flags;
  }
  return launch_job(c, j, s);
}

// ---------------------------------------------------------------------------
// Multipass: pass k = m..1 runs with interSize, searchMargin and step times k
// (src/piv_parameters.jl:85-88); every pass after the first shifts each search
// region by the previous pass's field at the IA (mqpiv_multipass.hip) and adds
// the shift to the displacement it measures; SN comes from the last pass
// (multi_quickPIV.jl:86).  `base` carries the final pass's buffers.
static int run_multipass(mqpiv_ctx* c, const mqpiv_params* p, int32_t m, const Job& base, hipStream_t s) {
  const int nd = p->ndims;
  std::vector<Job> js(m + 1);
  size_t vf_max = 0, shift_max = 0;
  for (int k = m; k >= 2; --k) {
    mqpiv_params q = *p;
    for (int d = 0; d < nd; ++d) {
      q.inter_size[d] *= k;
      q.search_margin[d] *= k;
      q.step[d] *= k;
    }
    q.compute_sn = 0;
    // buffers are placed below; a non-null VF placeholder passes validation
    int rc = build_job(c, &q, base.dims, base.in_dtype, base.img1, base.img2, base.mask, nullptr, MQPIV_F64,
                       (void*)&js, nullptr, &js[k]);
    if (rc) return rc;
    vf_max = std::max(vf_max, (size_t)js[k].n_ia * nd * sizeof(double));
  }
  js[1] = base;
  for (int k = m - 1; k >= 1; --k) shift_max = std::max(shift_max, (size_t)js[k].n_ia * 3 * sizeof(int32_t));
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t need = 2 * al(vf_max) + al(shift_max);
  if (need) {
    int rc = ensure(c, &c->d_mp, &c->mp_cap, need, s);
    if (rc) return rc;
  }
  char* slot[2] = {(char*)c->d_mp, (char*)c->d_mp + al(vf_max)};
  int32_t* shift = (int32_t*)((char*)c->d_mp + 2 * al(vf_max));
  const Job* prev = nullptr;
  for (int k = m; k >= 1; --k) {
    Job& j = js[k];
    if (k >= 2) j.vf = slot[k & 1];
    j.shift = nullptr;
    if (prev) {
      hipError_t e = launch_coarse_to_shift((const double*)prev->vf, prev->vf_n, prev->is, prev->st, j.vf_n, j.is,
                                            j.st, nd, j.n_ia, shift, s);
      if (e != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "multipass shift launch failed: %s", hipGetErrorString(e));
      j.shift = shift;
    }
    int rc = launch_job(c, j, s);
    if (rc) return rc;
    prev = j.n_ia > 0 ? &j : nullptr;
  }
  return MQPIV_OK;
}

static int check_multipass(mqpiv_ctx* c, const mqpiv_params* p, int32_t m) {
  if (m < 1 || m > 64) return fail(c, MQPIV_ERR_BAD_PARAMS, "multipass = %d (1..64)", m);
  if (m > 1 && p && p->precision_mode == MQPIV_SPLIT_MFMA)
    return fail(c, MQPIV_ERR_UNSUPPORTED, "the split-precision matrix-core form runs single passes");
  return MQPIV_OK;
}

extern "C" int mqpiv_piv_multipass(mqpiv_ctx* c, const mqpiv_params* p, int32_t multipass, const int64_t* dims,
                                   int32_t in_dtype, const void* img1, const void* img2, const void* mask,
                                   int32_t out_dtype, void* VF, void* SN) {
  if (!c) return MQPIV_ERR_NULL;
  int rc = check_multipass(c, p, multipass);
  if (rc) return rc;
  return run_staged(c, p, dims, in_dtype, img1, img2, mask, nullptr, out_dtype, VF, SN, nullptr,
                    [&](Job& j, hipStream_t s) { return run_multipass(c, p, multipass, j, s); });
}

extern "C" int mqpiv_piv_multipass_device(mqpiv_ctx* c, const mqpiv_params* p, int32_t multipass,
                                          const int64_t* dims, int32_t in_dtype, const void* d_img1,
                                          const void* d_img2, const void* d_mask, int32_t out_dtype, void* d_VF,
                                          void* d_SN, void* stream) {
  if (!c) return MQPIV_ERR_NULL;
  std::lock_guard<std::mutex> lk(c->mu);
  int rc = check_multipass(c, p, multipass);
  if (rc) return rc;
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "hipSetDevice failed");
  Job j;
  rc = build_job(c, p, dims, in_dtype, d_img1, d_img2, d_mask, nullptr, out_dtype, d_VF, d_SN, &j);
  if (rc) return rc;
  return run_multipass(c, p, multipass, j, stream ? (hipStream_t)stream : c->stream);
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

// ===========================================================================
// Batches

namespace {

struct BatchGroup {
  int fp64 = 0, L = 0, alg = 0;
  std::vector<int> members;  // indices into BatchPlan::jobs2d
  int64_t n_ia = 0;
  int ns0cap = 1, ns1cap = 1, ss0cap = 1;
  int Lr = 0;            // precision 3: the FP64 transform length of every member
  size_t table_off = 0;  // byte offset of [Job x m][int64 prefix x (m+1)]
};

struct BatchPlan {
  std::vector<Job> jobs2d;
  std::vector<size_t> sat_off;  // per 2-D job, offset of its integral arrays (or SIZE_MAX)
  std::vector<int> nsat;        // integral arrays a 2-D job builds (0 when it reads another job's)
  std::vector<int> need;        // integral arrays a 2-D job reads
  std::vector<int> sat_src;     // the job whose arrays it reads (itself when it builds them)
  std::vector<int> req2d, req3d;  // request index of each 2-D / 3-D job
  std::vector<Job> jobs3d;
  std::vector<BatchGroup> groups;
  size_t sat_total = 0, sat_tmp = 0, table_bytes = 0;
  size_t gz_bytes = 0;  // global transform buffers (shared by the groups, which run in turn)
};

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

static int plan_batch(mqpiv_ctx* c, const mqpiv_request* r, int n, BatchPlan& P) {
  if (n < 0 || (n > 0 && !r)) return fail(c, MQPIV_ERR_NULL, "requests is NULL");
  std::map<std::tuple<int, int, int, int>, int> key2g;
  // jobs that build integral arrays, by second image (pointer, type, size):
  // a later job of the same image reads the arrays of one that builds at
  // least as many (power 2, then power 1) instead of building its own
  std::map<std::tuple<const void*, int, int64_t, int64_t>, std::vector<int>> img2src;
  for (int i = 0; i < n; ++i) {
    Job j;
    int rc = build_job(c, &r[i].params, r[i].dims, r[i].in_dtype, r[i].img1, r[i].img2, r[i].mask, r[i].skip_mask,
                       r[i].out_dtype, r[i].VF, r[i].SN, &j);
    if (rc) return fail(c, rc, "request %d: %s", i, c->err.c_str());
    j.peaks = r[i].peaks;
    if (j.n_ia == 0) continue;
    if (j.ndims == 3) {
      P.jobs3d.push_back(j);
      P.req3d.push_back(i);
      continue;
    }
    int fp64 = 0, L = 0;
    rc = choose_2d(c, j, &fp64, &L);
    if (rc) return fail(c, rc, "request %d: %s", i, c->err.c_str());
    const int ns = sat_count(j.alg, 2, fp64);
    const int idx = (int)P.jobs2d.size();
    P.jobs2d.push_back(j);
    P.req2d.push_back(i);
    int src = idx;
    if (ns) {
      auto& srcs = img2src[std::make_tuple(j.img2, j.in_dtype, j.dims[0], j.dims[1])];
      for (int o : srcs)
        if (P.nsat[o] >= ns) src = o;
      if (src == idx) srcs.push_back(idx);
    }
    P.need.push_back(ns);
    P.sat_src.push_back(src);
    P.nsat.push_back(src == idx ? ns : 0);
    if (ns && src == idx) {
      P.sat_off.push_back(P.sat_total);
      P.sat_total += ns * al256(sat_bytes(j.dims, 2));
      P.sat_tmp = std::max(P.sat_tmp, al256(sat_tmp_bytes(j.dims, 2)));
    } else {
      P.sat_off.push_back(SIZE_MAX);
    }
    // precision 3 groups by its FP64 length too: a recomputed IA then gets the
    // values of the same call made alone
    const int Lr = fp64 >= 3 ? refine_len(j) : 0;
    auto key = std::make_tuple(fp64, L, (int)j.alg, Lr);
    auto it = key2g.find(key);
    int g;
    if (it == key2g.end()) {
      g = (int)P.groups.size();
      key2g[key] = g;
      BatchGroup bg;
      bg.fp64 = fp64;
      bg.L = L;
      bg.alg = j.alg;
      P.groups.push_back(bg);
    } else {
      g = it->second;
    }
    BatchGroup& bg = P.groups[g];
    bg.members.push_back(idx);
    bg.n_ia += j.n_ia;
    if (bg.n_ia >= ((int64_t)1 << 31))
      return fail(c, MQPIV_ERR_UNSUPPORTED, "request %d: a batch group holds 2^31 IAs or more", i);
    bg.ns0cap = std::max(bg.ns0cap, 2 * j.sm[0] + 1);
    bg.ns1cap = std::max(bg.ns1cap, 2 * j.sm[1] + 1);
    bg.ss0cap = std::max(bg.ss0cap, j.is[0] + 2 * j.sm[0]);
    bg.Lr = Lr;
  }
  for (auto& g : P.groups) {
    if (lds_2d(g.fp64, g.L, g.alg, g.ns0cap, g.ns1cap, g.ss0cap) > 160 * 1024 ||
        (g.fp64 >= 3 && lds_2d(1, g.Lr, g.alg, g.ns0cap, g.ns1cap, g.ss0cap) > 160 * 1024))
      return fail(c, MQPIV_ERR_UNSUPPORTED, "batch group (L=%d) does not fit in LDS", g.L);
    g.table_off = P.table_bytes;
    P.gz_bytes = std::max(P.gz_bytes, gz_scratch_2d(g.fp64, g.L, g.n_ia));
    P.table_bytes += al256(g.members.size() * sizeof(Job) + (g.members.size() + 1) * sizeof(int64_t));
  }
  return MQPIV_OK;
}

// Fills the job tables (host image of the device tables at d_table) with the
// integral-array pointers: inside sat_base, or in the building request's own
// workspace ws[request] when ws is given.
static void fill_tables(BatchPlan& P, char* sat_base, char* host_table, void* const* ws = nullptr) {
  for (size_t i = 0; i < P.jobs2d.size(); ++i) {
    Job& j = P.jobs2d[i];
    if (P.need[i]) {
      const int o = P.sat_src[i];
      const size_t one = al256(sat_bytes(j.dims, 2));
      char* b = ws ? (char*)ws[P.req2d[o]] : sat_base + P.sat_off[o];
      j.sat2 = (const double*)b;
      j.sat1 = P.need[i] >= 2 ? (const double*)(b + one) : nullptr;
    }
  }
  for (auto& g : P.groups) {
    Job* jt = (Job*)(host_table + g.table_off);
    int64_t* pf = (int64_t*)(host_table + g.table_off + g.members.size() * sizeof(Job));
    int64_t acc = 0;
    for (size_t m = 0; m < g.members.size(); ++m) {
      jt[m] = P.jobs2d[g.members[m]];
      pf[m] = acc;
      acc += jt[m].n_ia;
    }
    pf[g.members.size()] = acc;
  }
}

// Integral arrays + one launch per group (device job tables already in place).
// The prefix-scan temporary follows the arrays: at sat_base + sat_total, or in
// each request's workspace after its own arrays.
static int enqueue_2d(mqpiv_ctx* c, BatchPlan& P, char* sat_base, char* d_table, char* gz, hipStream_t s,
                      bool own_ws = false) {
  for (size_t i = 0; i < P.jobs2d.size(); ++i) {
    const Job& j = P.jobs2d[i];
    if (!P.nsat[i]) continue;
    double* tmp = own_ws ? (double*)((char*)j.sat2 + P.nsat[i] * al256(sat_bytes(j.dims, 2)))
                         : (double*)(sat_base + P.sat_total);
    hipError_t e = j.sat1 ? build_sat_pair(j.img2, j.in_dtype, j.dims, 2, (double*)j.sat2, (double*)j.sat1, tmp, s)
                          : build_sat(j.img2, j.in_dtype, 2, j.dims, 2, (double*)j.sat2, tmp, s);
    if (e != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "integral arrays: %s", hipGetErrorString(e));
  }
  for (auto& g : P.groups) {
    // a group of one pair passes its job in the kernel arguments (read through
    // the constant cache) instead of the device table, whose record every
    // workgroup would load before it can address its windows
    const bool one = g.members.size() == 1;
    const Job* dj = one ? nullptr : (const Job*)(d_table + g.table_off);
    const int64_t* dp = one ? nullptr : (const int64_t*)(d_table + g.table_off + g.members.size() * sizeof(Job));
    hipError_t e;
    {
      KernelTimer kt(c, s);
      e = launch_piv2d(g.fp64, g.L, g.alg, P.jobs2d[g.members[0]], dj, dp,
                       (int)g.members.size(), g.n_ia, g.ns0cap, g.ns1cap, g.ss0cap, gz, s, g.Lr);
    }
    if (e != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "batch launch failed: %s", hipGetErrorString(e));
  }
  return MQPIV_OK;
}

int batch_device_ws(mqpiv_ctx* c, const mqpiv_request* reqs, int32_t n, void* stream, void* const* ws) {
  if (!c) return MQPIV_ERR_NULL;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "hipSetDevice failed");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  BatchPlan P;
  int rc = plan_batch(c, reqs, n, P);
  if (rc) return rc;
  if (!P.groups.empty()) {
    if (!ws) {
      rc = ensure(c, &c->d_bsat, &c->bsat_cap, P.sat_total + P.sat_tmp + 256, s);
      if (rc) return rc;
    }
    if (P.gz_bytes) {
      rc = ensure_gz(c, P.gz_bytes, s);
      if (rc) return rc;
    }
    // Groups of one pair take their job in the kernel arguments; only groups of
    // several read the device job tables, and only then are they copied.
    bool tables = false;
    for (const auto& g : P.groups) tables |= g.members.size() > 1;
    if (tables) {
      rc = ensure(c, &c->d_table, &c->table_cap, std::max(P.table_bytes + 256, kMinTableBytes), s);
      if (rc) return rc;
      // pinned staging for the job tables; the previous batch's copy must be done
      if (c->table_done) hipEventSynchronize(c->table_done);
      if (P.table_bytes > c->h_table_cap) {
        const size_t want = std::max(P.table_bytes, kMinTableBytes);
        if (c->h_table) hipHostFree(c->h_table);
        c->h_table = nullptr;
        c->h_table_cap = 0;
        HIP_OR_FAIL(c, hipHostMalloc(&c->h_table, want, hipHostMallocDefault));
        c->h_table_cap = want;
      }
      if (!c->table_done) HIP_OR_FAIL(c, hipEventCreateWithFlags(&c->table_done, hipEventDisableTiming));
      fill_tables(P, (char*)c->d_bsat, (char*)c->h_table, ws);
      HIP_OR_FAIL(c, hipMemcpyAsync(c->d_table, c->h_table, P.table_bytes, hipMemcpyHostToDevice, s));
      HIP_OR_FAIL(c, hipEventRecord(c->table_done, s));
    } else {
      std::vector<char> host(P.table_bytes + 1);  // the jobs' integral-array pointers only
      fill_tables(P, (char*)c->d_bsat, host.data(), ws);
    }
    rc = enqueue_2d(c, P, (char*)c->d_bsat, tables ? (char*)c->d_table : nullptr, (char*)c->d_gz, s, ws != nullptr);
    if (rc) return rc;
  }
  for (size_t k = 0; k < P.jobs3d.size(); ++k) {
    rc = launch_job(c, P.jobs3d[k], s, ws ? (char*)ws[P.req3d[k]] : nullptr);
    if (rc) return rc;
  }
  return MQPIV_OK;
}

extern "C" int mqpiv_piv_batch_device(mqpiv_ctx* c, const mqpiv_request* reqs, int32_t n, void* stream) {
  return batch_device_ws(c, reqs, n, stream, nullptr);
}

struct mqpiv_graph {
  mqpiv_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void* d_mem = nullptr;  // integral arrays + job tables owned by the graph
};

extern "C" int mqpiv_graph_create(mqpiv_ctx* c, const mqpiv_request* reqs, int32_t n, mqpiv_graph** out) {
  if (!c || !out) return MQPIV_ERR_NULL;
  *out = nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MQPIV_ERR_DEVICE, "hipSetDevice failed");
  BatchPlan P;
  int rc = plan_batch(c, reqs, n, P);
  if (rc) return rc;
  if (!P.jobs3d.empty()) return fail(c, MQPIV_ERR_UNSUPPORTED, "graph capture covers 2-D pairs only");
  mqpiv_graph* g = new mqpiv_graph();
  g->ctx = c;
  const size_t sat_part = al256(P.sat_total + P.sat_tmp + 256);
  const size_t table_part = al256(P.table_bytes + 256);
  hipError_t e = hipMalloc(&g->d_mem, sat_part + table_part + P.gz_bytes);
  if (e != hipSuccess) {
    delete g;
    return fail(c, MQPIV_ERR_OOM, "graph memory: %s", hipGetErrorString(e));
  }
  char* sat_base = (char*)g->d_mem;
  char* d_table = sat_base + sat_part;
  char* gz = P.gz_bytes ? d_table + table_part : nullptr;
  std::vector<char> host(P.table_bytes + 1);
  fill_tables(P, sat_base, host.data());
  e = hipMemcpy(d_table, host.data(), P.table_bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess && gz) e = hipMemset(gz, 0, kGzFlagBytes);  // slot flags, before capture
  if (e == hipSuccess) e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    hipFree(g->d_mem);
    delete g;
    return fail(c, MQPIV_ERR_DEVICE, "graph setup: %s", hipGetErrorString(e));
  }
  const bool prof = c->profiling;
  c->profiling = false;
  rc = enqueue_2d(c, P, sat_base, d_table, gz, c->stream);
  c->profiling = prof;
  e = hipStreamEndCapture(c->stream, &g->graph);
  if (rc == MQPIV_OK && e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (rc != MQPIV_OK || e != hipSuccess) {
    if (g->graph) hipGraphDestroy(g->graph);
    hipFree(g->d_mem);
    delete g;
    return rc ? rc : fail(c, MQPIV_ERR_DEVICE, "graph capture: %s", hipGetErrorString(e));
  }
  *out = g;
  return MQPIV_OK;
}

extern "C" int mqpiv_graph_launch(mqpiv_graph* g, void* stream) {
  if (!g) return MQPIV_ERR_NULL;
  hipStream_t s = stream ? (hipStream_t)stream : g->ctx->stream;
  hipError_t e = hipGraphLaunch(g->exec, s);
  if (e != hipSuccess) return fail(g->ctx, MQPIV_ERR_DEVICE, "hipGraphLaunch: %s", hipGetErrorString(e));
  return MQPIV_OK;
}

extern "C" int mqpiv_graph_destroy(mqpiv_graph* g) {
  if (!g) return MQPIV_ERR_NULL;
  hipDeviceSynchronize();
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  if (g->d_mem) hipFree(g->d_mem);
  delete g;
  return MQPIV_OK;
}
