// mqpiv_internal.h -- descriptors shared by the host runtime and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <set>
#include <type_traits>
#include <utility>

namespace mqpiv {

enum Alg : int32_t { ALG_FFTCC = 0, ALG_ZNCC = 1, ALG_NSQECC = 2, ALG_MASK_NSQECC = 3 };
// Job::precision_mode of the split-precision matrix-core form (MQPIV_SPLIT_MFMA)
constexpr int32_t kPrecSplitMfma = 1;
enum DType : int32_t { DT_F32 = 0, DT_F64 = 1 };

// One frame pair (a "job") as the kernels see it.  Geometry follows the
// reference's per-IA coordinates (src/utils/piv_ops.jl:38-85), 0-based here.
struct Job {
  const void* img1;
  const void* img2;
  const void* mask;        // masked NSQECC only
  const uint8_t* skip;     // optional host-computed skip flags (one per IA)
  void* vf;                // N x nVF, component fastest
  void* sn;                // nVF, or nullptr
  int32_t* peaks;          // optional: N int32 per IA, 1-based integer peak (0 = none)
  const int32_t* shift;    // optional: 3 int32 per IA, search-region shift (multipass)
  const double* sat1;      // FP64 integral array of img2 (sum of pixels), or nullptr
  const double* sat2;      // FP64 integral array of img2^2, or nullptr
  const double* sat1f;     // 3-D ZNCC means: FP64 integral array of img1 (with sat1 the
                           // one of img2), or, with sat1 NULL, the per-IA sums (img1
                           // over the IA at [ia], img2 over the search region at
                           // [n_ia + ia]); else nullptr
  int64_t dims[3];
  int32_t is[3], sm[3], st[3];
  int32_t vf_n[3];         // vector-field size per dim
  int64_t n_ia;
  int64_t ia_base;         // first global IA index of this job in a batch
  int32_t ndims;
  int32_t alg;
  int32_t in_dtype;
  int32_t out_dtype;
  int32_t compute_sn;
  int32_t has_threshold;   // unmasked: threshold >= 0; masked: threshold > 0
  int32_t has_mask_threshold;
  int32_t precision_mode;  // launch choice (MQPIV_PLAIN_FP32 / SPLIT_MFMA / FP64 / AUTO)
  double threshold;
  double mask_threshold;
};

// Launch-wide description of a batch of jobs sharing one kernel instance.
struct Batch {
  const Job* jobs;         // device array
  int32_t n_jobs;
  int32_t pad_;
  int64_t n_ia_total;
};

// 2-D transform lengths whose buffer does not fit in LDS: FP64 at 128 needs
// 128 x 129 x 16 B = 258 KB, more than a CU's 160 KB.  Its buffer lives in
// global memory instead: one of gz_slots_2d buffers of gz_stride_2d bytes,
// which a block claims for its IA (an int32 flag per slot, 0 = free, in the
// kGzFlagBytes before the buffers) and frees at its end; the L2 holds the
// buffer while the block works on it.  At most kGzBlocksPerCu blocks of such a
// kernel are resident per CU (registers), so that many slots per CU (or one
// per block, if fewer) leave every claiming block a free slot.
template <int L, typename T>
constexpr bool gz_len() {
  return std::is_same<T, double>::value && L == 128;
}
template <int L, typename T>
constexpr size_t gz_stride_2d() {
  return ((size_t)L * (L + 1) * 2 * sizeof(T) + 4095) & ~(size_t)4095;
}
constexpr int kGzBlocksPerCu = 2;
constexpr int kGzMaxSlots = 1024;
constexpr size_t kGzFlagBytes = kGzMaxSlots * sizeof(int32_t);
constexpr uint32_t kGzProbeLimit = 1u << 24;  // probes before a block gives up (NaN outputs)

// Refine list of a split-precision ZNCC launch: the IAs whose results the FP64
// transforms recompute.  int32 count at [0], the global IA indices of the
// launch (the order in which the flagging kernel appended them) from
// [kRefineHead]; at most one entry per IA.
constexpr int kRefineHead = 64;
inline size_t refine_bytes(int64_t n_ia) { return ((size_t)(kRefineHead + n_ia) * sizeof(int32_t) + 255) & ~(size_t)255; }

// Compute units of the current device (cached per device).
inline int device_cus() {
  static std::mutex mu;
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  std::lock_guard<std::mutex> lock(mu);
  if (!cus[dev] && hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus[dev] = 256;
  return cus[dev] > 0 ? cus[dev] : 256;
}

// Global transform-buffer slots for a launch of n_ia IAs.
inline int gz_slots_2d(int64_t n_ia) {
  int64_t c = (int64_t)kGzBlocksPerCu * device_cus();
  if (c > kGzMaxSlots) c = kGzMaxSlots;
  return (int)(n_ia < c ? n_ia : c);
}

// Raises a kernel's dynamic-LDS limit to the CU's 160 KB on the current device,
// once per (kernel, device); safe to call from any thread before a launch.
inline void allow_full_lds(const void* fn) {
  static std::mutex mu;
  static std::set<std::pair<const void*, int>> done;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  std::lock_guard<std::mutex> lock(mu);
  if (done.insert({fn, dev}).second &&
      hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
    (void)hipGetLastError();  // a kernel with static LDS keeps its default; clear the error for the launch check
}

}  // namespace mqpiv
