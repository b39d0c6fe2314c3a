/*
 * mqpiv.h -- C ABI of libmqpiv, the MI355X-native cross-correlation PIV hot path.
 *
 * Plain pointers and sizes only.  Every image/volume buffer is column-major with
 * dim 1 fastest (exactly a Julia Array's memory), and the outputs use the
 * reference layout:
 *   VF : ndims x nVF_1 x nVF_2 [x nVF_3], component index fastest
 *        (src/utils/piv_ops.jl:25-31, allocate_outputs)
 *   SN : nVF_1 x nVF_2 [x nVF_3]
 * IAs skipped by the threshold test keep VF = 0 and SN = 0, as in the reference.
 *
 * Entry points and the reference interface each one replaces (files relative
 * to the reference repository root):
 *
 *   mqpiv_ctx_create   allocate_tmp_data(alg, pivparams, T) + FFTW planning
 *                      (src/crosscorrelation_algorithms/fftcc.jl:26-51,
 *                      zncc.jl:27-53, nsqecc.jl:30-56, mask_nsqecc.jl:30-57;
 *                      src/utils/fft_ops.jl:11-55).  Done once per process
 *                      instead of once per PIV call.
 *   mqpiv_ctx_destroy  destroy_fftw_plans (fftcc.jl:78-82, fft_ops.jl:72-83)
 *   mqpiv_vf_size      get_vectorfield_size (src/utils/piv_ops.jl:6-14)
 *   mqpiv_piv          PIV_CPU (src/multi_quickPIV.jl:40-96): for every IA,
 *                      prepare_inputs! + crosscorrelation! (the plugin quartet of
 *                      piv_parameters.jl:2-7) + gaussian_displacement
 *                      (piv_ops.jl:249-273) + compute_SN (piv_ops.jl:333-367),
 *                      host buffers in and out.  With corr_alg = MQPIV_MASK_NSQECC
 *                      and a mask, the masked drivers' intended semantics
 *                      (multi_quickPIV.jl:99-154).
 *   mqpiv_piv_device   the same on device-resident buffers, asynchronous on the
 *                      caller's HIP stream (no reference counterpart: it is the
 *                      resident-data form used by batching and serving).
 *   mqpiv_last_error   the message of the last failure on this context
 *                      (the reference raises AssertionError/MethodError instead).
 */
#ifndef MQPIV_H
#define MQPIV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define MQPIV_OK 0
#define MQPIV_ERR_SIZE_MISMATCH 1 /* inputs differ in size (reference: AssertionError) */
#define MQPIV_ERR_BAD_PARAMS 2    /* invalid parameter values                           */
#define MQPIV_ERR_UNSUPPORTED 3   /* valid in the reference but not built here          */
#define MQPIV_ERR_OOM 4           /* device allocation failed / over the HBM budget     */
#define MQPIV_ERR_DEVICE 5        /* a HIP runtime error                                */
#define MQPIV_ERR_NULL 6          /* a required pointer was NULL                         */

/* Correlation algorithms (src/piv_parameters.jl:2-7, parseCorr :14-23). */
#define MQPIV_FFTCC 0
#define MQPIV_ZNCC 1
#define MQPIV_NSQECC 2
#define MQPIV_MASK_NSQECC 3

/* Element types of image and output buffers. */
#define MQPIV_F32 0
#define MQPIV_F64 1

/* Precision modes (mqpiv_params.precision_mode). */
#define MQPIV_PLAIN_FP32 0 /* FP32 transforms; FP64 window sums, sum(F^2) and denominators    */
#define MQPIV_SPLIT_MFMA 1 /* 2-D DFTs as matrix-core products with hi + lo FP16 operands,
                              FP32 accumulation; FP64 sums and denominators as in mode 0.
                              FFTCC, ZNCC and NSQECC with interSize <= 32 and search size
                              <= 64, or interSize <= 64 and margin <= 16, windows up to
                              48 x 48 (single calls, batches, graphs, the server); no
                              multipass                                                       */
#define MQPIV_FP64 2       /* every step in FP64 (2-D search sizes up to 128)                    */
#define MQPIV_AUTO 3       /* FP64 where FP32-level transforms miss the parity bar (masked
                              NSQECC, and ZNCC when SN is requested); elsewhere the split-
                              precision form where it takes the job, FP32 transforms otherwise.
                              ZNCC with SN runs the split-precision form where it takes the
                              job, and the FP64 transforms recompute the IAs whose SN, peak
                              or subpixel fit it cannot settle to the FP64 results' bar
                              (form 3 of mqpiv_transform_form); 2-D masked NSQECC runs FP32
                              transforms with the same recomputation (form 4)                */

/* The fields of PIVParameters (src/piv_parameters.jl:29-46) that the device
 * needs, already resolved: step = interSize - overlap when the reference's
 * step is `nothing` (piv_parameters.jl:88).  Unused trailing dims are ignored. */
typedef struct mqpiv_params {
  int32_t corr_alg;
  int32_t ndims;             /* 2 or 3 */
  int64_t inter_size[3];
  int64_t search_margin[3];
  int64_t step[3];
  int32_t compute_sn;
  int32_t precision_mode;
  double threshold;          /* reference default -1.0 (no skipping) */
  double mask_threshold;     /* reference default -1.0 */
} mqpiv_params;

typedef struct mqpiv_ctx mqpiv_ctx;

/* Create a context on HIP device `device`.  hbm_budget_bytes = 0 means no limit
 * beyond the device's memory.  Returns MQPIV_OK and sets *out.  A context's
 * device buffers (integral arrays, job tables, refine lists, transform
 * buffers) serve every call made through it: calls through one context must
 * not overlap on the device (one stream per context, as the server's slots). */
int mqpiv_ctx_create(int device, uint64_t hbm_budget_bytes, mqpiv_ctx** out);
int mqpiv_ctx_destroy(mqpiv_ctx* ctx);
const char* mqpiv_last_error(const mqpiv_ctx* ctx);

/* nVF per dim and the IA count for an image of size dims[0..ndims-1]. */
int mqpiv_vf_size(const mqpiv_params* p, const int64_t* dims, int64_t* vf_out3, int64_t* n_ia);

/* The transform form a call with these parameters and dims runs: *form = 0
 * (FP32 transforms), 1 (FP64 transforms), 2 (the split-precision matrix-core
 * form), 3 (that form, then FP64 transforms over the IAs it flags; ZNCC with
 * SN in MQPIV_AUTO) or 4 (FP32 transforms, then FP64 transforms of the same
 * length over the IAs they flag; 2-D masked NSQECC in MQPIV_AUTO), *length =
 * the transform length (2-D; for the split form
 * 64, or 65 for IAs taken in 2 x 2 sub-blocks of length-64 products).  Host-only; no device
 * work.  Same errors as the call itself would report for an unsupported job.
 * No counterpart in the reference (its transforms are always FFTW at
 * good_size_real, src/utils/fft_ops.jl:145-198); it lets a caller see which
 * form the precision mode chose. */
int mqpiv_transform_form(const mqpiv_params* p, const int64_t* dims, int32_t* form, int32_t* length);

/* Allocation-free form of mqpiv_piv_device.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc or hipFree: the integral arrays of ZNCC and NSQECC live in a static
 * per-device array of MQPIV_NOALLOC_SCRATCH_BYTES (96 MiB: a 2048 x 2048 pair with
 * ZNCC fits), which each GPU holds from library load until the process ends;
 * the context must have been created on that device.  2-D pairs only, and not
 * FP64 transforms at search size 128 (MQPIV_ERR_UNSUPPORTED); a pair whose
 * integral arrays exceed the array is rejected with MQPIV_ERR_OOM.  Every call
 * on a device shares its array, which keeps the last call's contents: calls
 * that overlap on one device (other streams, other contexts) race.  Results
 * equal mqpiv_piv_device's. */
#define MQPIV_NOALLOC_SCRATCH_BYTES (96ull << 20)
int mqpiv_piv_device_noalloc(mqpiv_ctx* ctx, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                             const void* d_img1, const void* d_img2, const void* d_mask,
                             const uint8_t* d_skip_mask, int32_t out_dtype, void* d_VF, void* d_SN,
                             void* stream);

/* Host buffers.  img1/img2 (and mask) have in_dtype elements; VF/SN have
 * out_dtype elements and are fully written (SN may be NULL; it is required
 * when compute_sn is set).  skip_mask (nullable, one byte per IA) overrides the
 * threshold test -- used when the caller's filtFun is not the default maxval. */
int mqpiv_piv(mqpiv_ctx* ctx, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
              const void* img1, const void* img2, const void* mask, const uint8_t* skip_mask,
              int32_t out_dtype, void* VF, void* SN);

/* mqpiv_piv plus a diagnostic output: peaks (nullable) receives, per IA, the
 * ndims 1-based integer peak indices of first_fullovp_peak
 * (src/utils/common_ops.jl:98-124), 0 when no value beats 0 or the IA is skipped. */
int mqpiv_piv_ex(mqpiv_ctx* ctx, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                 const void* img1, const void* img2, const void* mask, const uint8_t* skip_mask,
                 int32_t out_dtype, void* VF, void* SN, int32_t* peaks);

/* Device buffers, enqueued on `stream` (a hipStream_t; NULL = the context's
 * stream) and not synchronised. */
int mqpiv_piv_device(mqpiv_ctx* ctx, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                     const void* d_img1, const void* d_img2, const void* d_mask,
                     const uint8_t* d_skip_mask, int32_t out_dtype, void* d_VF, void* d_SN,
                     void* stream);

/* PIV(img1, img2, mask_VF, mask_IA, pivparams) (src/multi_quickPIV.jl:155-213):
 * masked NSQECC (corr_alg must be MQPIV_MASK_NSQECC, 2-D) correlating with
 * mask_IA, while the skip filter of skip_inter_region (src/utils/piv_ops.jl:
 * 148-157, filtFun = mask_filtFun = maxval) reads image 1 and mask_VF.  The skip
 * flags are computed on the device.  Host buffers (mqpiv_piv_masks; peaks as in
 * mqpiv_piv_ex, nullable) or device buffers enqueued on `stream`
 * (mqpiv_piv_masks_device; NULL = the context's stream).  With one mask passed
 * as both, results equal mqpiv_piv's. */
int mqpiv_piv_masks(mqpiv_ctx* ctx, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                    const void* img1, const void* img2, const void* mask_vf, const void* mask_ia,
                    int32_t out_dtype, void* VF, void* SN, int32_t* peaks);
int mqpiv_piv_masks_device(mqpiv_ctx* ctx, const mqpiv_params* p, const int64_t* dims, int32_t in_dtype,
                           const void* d_img1, const void* d_img2, const void* d_mask_vf,
                           const void* d_mask_ia, int32_t out_dtype, void* d_VF, void* d_SN, void* stream);

/* Multipass PIV: the coarse-to-fine loop `for scale in multipass:-1:1` of
 * PIV_CPU (src/multi_quickPIV.jl:57-88), with the search-region offset hook of
 * get_search_coordinates(..., displacement) (src/utils/piv_ops.jl:59-76) wired
 * as piv_ops.jl:160-173 describes; the reference leaves both unfinished (its
 * passes ignore the offset, and scale > 1 only rescales the peak window).
 * Pass k = multipass..1 runs with interSize, searchMargin and step times k
 * (piv_parameters.jl:85-88).  Each pass after the first shifts every search
 * region by d = the previous pass's vector field interpolated (multilinear, in
 * vector-field index space) at the IA's centre and rounded to the nearest
 * integer (ties to even; a NaN input gives 0), clamped so that the IA moved by d
 * stays inside the image.  The shifted search region is the one the reference
 * builds for an IA at TLF + d (its SR_TLF_off of piv_ops.jl:72 included); the
 * IA stays at TLF, and the pass reports d + its measured displacement.  VF and
 * SN (computeSN: last pass only) have the multipass = 1 shapes.  Every pass must fit a transform
 * (2-D search size <= 128, 3-D <= 64; MQPIV_ERR_UNSUPPORTED otherwise).  The
 * threshold test runs on each pass's own IAs; there is no skip mask.
 * multipass = 1 equals mqpiv_piv. */
int mqpiv_piv_multipass(mqpiv_ctx* ctx, const mqpiv_params* p, int32_t multipass, const int64_t* dims,
                        int32_t in_dtype, const void* img1, const void* img2, const void* mask,
                        int32_t out_dtype, void* VF, void* SN);
/* The same over device buffers, enqueued on `stream` (NULL = the context's stream). */
int mqpiv_piv_multipass_device(mqpiv_ctx* ctx, const mqpiv_params* p, int32_t multipass, const int64_t* dims,
                               int32_t in_dtype, const void* d_img1, const void* d_img2, const void* d_mask,
                               int32_t out_dtype, void* d_VF, void* d_SN, void* stream);

/* One frame pair of a batch (device buffers). */
typedef struct mqpiv_request {
  mqpiv_params params;
  int64_t dims[3];
  int32_t in_dtype;
  int32_t out_dtype;
  const void* img1;
  const void* img2;
  const void* mask;          /* masked NSQECC only */
  const uint8_t* skip_mask;  /* nullable */
  void* VF;
  void* SN;                  /* nullable unless compute_sn */
  int32_t* peaks;            /* nullable */
} mqpiv_request;

/* A batch of frame pairs of any sizes and parameters, device buffers (no
 * reference counterpart: the reference processes one pair per PIV call).  2-D
 * pairs that share algorithm, precision and transform length run in ONE kernel
 * launch over all their IAs; 3-D pairs run one after another.  Enqueued on
 * `stream` (NULL = the context's), not synchronised.  Results per pair are
 * identical to separate calls: every IA is computed independently. */
int mqpiv_piv_batch_device(mqpiv_ctx* ctx, const mqpiv_request* reqs, int32_t n, void* stream);

/* The same 2-D batch captured once into a HIP graph.  The buffers named in reqs
 * must stay valid for the graph's lifetime; every launch recomputes from their
 * current contents (the graph owns its integral-array and job-table memory). */
typedef struct mqpiv_graph mqpiv_graph;
int mqpiv_graph_create(mqpiv_ctx* ctx, const mqpiv_request* reqs, int32_t n, mqpiv_graph** out);
int mqpiv_graph_launch(mqpiv_graph* g, void* stream);
int mqpiv_graph_destroy(mqpiv_graph* g);

/* Measurement hooks (no reference counterpart).  Between profile_begin and
 * profile_end every correlation-kernel launch is bracketed by HIP events on the
 * stream it runs on; profile_end synchronises and returns the durations in
 * launch order (milliseconds; *n = count, at most cap are written). */
int mqpiv_profile_begin(mqpiv_ctx* ctx);
int mqpiv_profile_end(mqpiv_ctx* ctx, double* kernel_ms, int32_t cap, int32_t* n);

/* Continuous-batching request server (no reference counterpart: the reference
 * runs one PIV call at a time on the caller's thread).  A worker thread admits
 * submitted pairs, oldest first, into one of n_slots execution slots as soon
 * as a slot is free and the pair's device footprint fits in the
 * hbm_budget_bytes pool (up to max_batch pairs per slot).  The device runs the
 * pairs in submission order: a large pair (a 3-D volume, or a 2-D pair of more
 * than 32,768 IAs) runs alone, after the batches admitted before it and before
 * those admitted after it; batches of consecutive small pairs of one
 * dimensionality may run beside each other.  The order is kept on the device
 * (a slot's stream waits for the slots it follows), so a batch starts as soon
 * as the ones ahead of it end.
 * The footprint is the pair's FP64 integral arrays plus, for host requests, its
 * staged inputs and outputs, each rounded to 256 B; every slot also keeps its
 * own context (3-D transform scratch, at most MQPIV_3D_CHUNK_BYTES, 2 GiB by
 * default) outside the pool.  Each slot copies its host pairs in, runs all its pairs as one
 * batched call (as mqpiv_piv_batch_device) on its own stream, and copies
 * VF/SN/peaks back, so batches of different slots overlap on the device.
 *   mqpiv_server_submit: HOST pointers, valid until mqpiv_server_wait reports
 *     completion (pinned memory lets copies overlap kernels);
 *   mqpiv_server_submit_device: DEVICE pointers (frames resident in HBM); the
 *     outputs are in device memory at completion.
 * skip_mask is not accepted (MQPIV_ERR_UNSUPPORTED); a pair larger than the
 * whole pool is rejected with MQPIV_ERR_OOM at submit.  Results are identical
 * to a lone mqpiv_piv call: every IA is computed independently. */
typedef struct mqpiv_server mqpiv_server;
typedef struct {
  uint64_t completed; /* pairs finished OK since create / reset_stats */
  int64_t ia_done;
  uint64_t pending;   /* submitted, not yet admitted */
  double p50_ms, p99_ms, max_ms; /* submit -> results on host */
  double pairs_per_s, ia_per_s;  /* over first submit .. last completion */
} mqpiv_server_stats;
int mqpiv_server_create(int device, uint64_t hbm_budget_bytes, int32_t n_slots, int32_t max_batch,
                        mqpiv_server** out);
int mqpiv_server_submit(mqpiv_server* sv, const mqpiv_request* req, uint64_t* ticket);
int mqpiv_server_submit_device(mqpiv_server* sv, const mqpiv_request* req, uint64_t* ticket);
/* Waits up to timeout_s (< 0: forever); *status = the pair's MQPIV_* code, or -1
 * if still running (the ticket then stays valid). */
int mqpiv_server_wait(mqpiv_server* sv, uint64_t ticket, double timeout_s, int32_t* status);
int mqpiv_server_stats_get(mqpiv_server* sv, mqpiv_server_stats* st);
/* Message of the last failed batch (a batch fails as a whole: every pair in it
 * reports the same status).  Valid until the next call on this thread. */
const char* mqpiv_server_last_error(mqpiv_server* sv);
int mqpiv_server_reset_stats(mqpiv_server* sv);
int mqpiv_server_destroy(mqpiv_server* sv);

/* ---------------------------------------------------------------------------
 * Vector-field post-processing on the device (src/postprocessing/postprocessing.jl).
 * All buffers are device memory in the reference layout: VF is ndims x nVF...
 * (component fastest), per-vector maps (magnitudes, SN, masks, collectiveness)
 * are nVF... with dim 1 fastest.  VF and maps have the element type desc->dtype
 * (the PIV output precision); masks are one byte per vector (0/1).  Radii are
 * per dim (the reference's toDims3(radius)[1:N]); boxes are clipped at the
 * field border (localOp, postprocessing.jl:56-101).  Enqueued on `stream` (NULL
 * = the context's), not synchronised; they share the context's scratch, so one
 * stream at a time per context.  Outputs must not alias inputs. */
typedef struct mqpiv_vf_desc {
  int32_t ndims;   /* 2 or 3: components per vector = grid dims */
  int32_t dtype;   /* MQPIV_F32 / MQPIV_F64 */
  int64_t size[3]; /* nVF per dim (size[2] ignored in 2-D) */
} mqpiv_vf_desc;

#define MQPIV_CMP_GT 0       /* mask = map > th (TH_masking's default cmp)   */
#define MQPIV_CMP_LT 1       /* mask = map < th (remove_low_SN)              */
#define MQPIV_REPLACE_ZERO 0 /* zeroReplace                                  */
#define MQPIV_REPLACE_MEDIAN 1 /* medianReplace, and meanReplace (its body is the median) */

/* magnitudes(VF): M = sqrt(sum(VF[:,i].^2))   (postprocessing.jl:178-194) */
int mqpiv_pp_magnitudes(mqpiv_ctx* ctx, const mqpiv_vf_desc* desc, const void* VF, void* M, void* stream);
/* STD_masking(n, map): |map - mean| > n * std, corrected std (:598-610) */
int mqpiv_pp_std_mask(mqpiv_ctx* ctx, const mqpiv_vf_desc* desc, const void* map, double n_std, uint8_t* mask,
                      void* stream);
/* TH_masking(th, map; cmp) (:640-663) */
int mqpiv_pp_th_mask(mqpiv_ctx* ctx, const mqpiv_vf_desc* desc, const void* map, double th, int32_t cmp,
                     uint8_t* mask, void* stream);
/* The replacement step of STD_filtering / TH_filtering / remove_low_SN
 * (:539-573, 618-634, 672-703, 712-732): out = VF, with every flagged vector
 * replaced by the zero vector or by the per-component median of the clipped
 * box (radius) in VF.  Median: NaN if the box holds a NaN; even counts give
 * a/2 + b/2 of the central pair (Statistics.median). */
int mqpiv_pp_replace(mqpiv_ctx* ctx, const mqpiv_vf_desc* desc, const void* VF, const uint8_t* mask, int32_t how,
                     const int64_t* radius, void* out, void* stream);
/* average(radius, VF): per-component box mean (:359-381) */
int mqpiv_pp_average(mqpiv_ctx* ctx, const mqpiv_vf_desc* desc, const void* VF, const int64_t* radius, void* out,
                     void* stream);
/* collectiveness(radius, VF; max_angle) (:294-331): Float32 fraction of the box
 * whose normalised dot product with the vector exceeds min_dot = cosd(max_angle). */
int mqpiv_pp_collectiveness(mqpiv_ctx* ctx, const mqpiv_vf_desc* desc, const void* VF, const int64_t* radius,
                            double min_dot, float* out, void* stream);
/* similarityAveraging(radius, VF; max_angle, normalize, scale_by_similarity,
 * similarity_exponent) (:478-523), min_dot = cosd(max_angle); exponent >= 0. */
int mqpiv_pp_similarity_averaging(mqpiv_ctx* ctx, const mqpiv_vf_desc* desc, const void* VF, const int64_t* radius,
                                  double min_dot, int32_t normalize, int32_t scale_by_similarity,
                                  int32_t similarity_exponent, void* out, void* stream);
/* local_magnitude_mask(ratio, VF; radius) (:202-218): out = VF with vectors whose
 * magnitude / mean magnitude of the rest of the box is not < ratio set to
 * (signed) zero; removed = 1 for those. */
int mqpiv_pp_local_magnitude_mask(mqpiv_ctx* ctx, const mqpiv_vf_desc* desc, const void* VF, double ratio,
                                  const int64_t* radius, void* out, uint8_t* removed, void* stream);

/* Library build identification (for logs). */
const char* mqpiv_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MQPIV_H */
