// Library-internal context state shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mqpiv.h"

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
  // batches: integral arrays of all pairs, device job tables, pinned staging
  void* d_bsat = nullptr;
  size_t bsat_cap = 0;
  void* d_table = nullptr;
  size_t table_cap = 0;
  void* h_table = nullptr;
  size_t h_table_cap = 0;
  hipEvent_t table_done = nullptr;
  // post-processing filters: reduction partials + a magnitude map
  void* d_pp = nullptr;
  size_t pp_cap = 0;
  // global transform buffers of the 2-D lengths that do not fit in LDS
  void* d_gz = nullptr;
  size_t gz_cap = 0;
  // multipass: two coarse vector fields (FP64) and one pass's search-region shifts
  void* d_mp = nullptr;
  size_t mp_cap = 0;
  // double-mask calls: per-IA skip flags computed on the device
  void* d_flags = nullptr;
  size_t flags_cap = 0;
  // measurement hooks
  bool profiling = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};

// Brackets one correlation-kernel launch with events when profiling is on.
// The events only time the launch (mqpiv_profile_end waits on them before
// reading the interval), so they skip the system-scope fence: no cache
// writeback and invalidation between the timed kernels.
struct KernelTimer {
  mqpiv_ctx* c;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  KernelTimer(mqpiv_ctx* c_, hipStream_t s_) : c(c_), s(s_) {
    if (c->profiling && hipEventCreateWithFlags(&a, hipEventDisableSystemFence) == hipSuccess &&
        hipEventCreateWithFlags(&b, hipEventDisableSystemFence) == hipSuccess)
      hipEventRecord(a, s);
  }
  ~KernelTimer() {
    if (a && b) {
      hipEventRecord(b, s);
      c->events.emplace_back(a, b);
    }
  }
};

// Records a failure message on the context and returns code.
int fail(mqpiv_ctx* c, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));

// Grow-only device buffer owned by the context (synchronises s before a regrow).
int ensure(mqpiv_ctx* c, void** buf, size_t* cap, size_t need, hipStream_t s);

#define HIP_OR_FAIL(ctx, expr)                                                              \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(ctx, e_ == hipErrorOutOfMemory ? MQPIV_ERR_OOM : MQPIV_ERR_DEVICE,        \
                  "%s failed: %s", #expr, hipGetErrorString(e_));                           \
  } while (0)

