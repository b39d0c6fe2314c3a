// mqpiv_server_core.h -- the request server's bookkeeping: the device-pool
// allocator, request records, and the completion and retirement of a slot's
// requests.  Free of HIP calls, so that it also builds for a CPU test.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#include <iterator>
#include <map>
#include <string>
#include <vector>

#include "../../include/mqpiv.h"

namespace mqpiv_srv {

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// First-fit allocator over one device region, with coalescing on free.
class Pool {
 public:
  void init(size_t bytes) {
    free_.clear();
    free_[0] = bytes;
    cap_ = bytes;
    used_ = 0;
  }
  size_t alloc(size_t n) {
    n = al256(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        size_t off = it->first, sz = it->second;
        free_.erase(it);
        if (sz > n) free_[off + n] = sz - n;
        used_ += n;
        return off;
      }
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t n) {
    n = al256(n);
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t capacity() const { return cap_; }
  size_t used() const { return used_; }

 private:
  std::map<size_t, size_t> free_;
  size_t cap_ = 0, used_ = 0;
};

struct Req {
  uint64_t id = 0;
  mqpiv_request req{};  // the client's buffers: host, or device when on_device
  bool on_device = false;
  int64_t n_ia = 0;
  size_t npx = 0, esz = 0, osz = 0;
  size_t io_bytes = 0, ws_bytes = 0;  // staged inputs/outputs (host requests) + integral arrays
  size_t d_off = SIZE_MAX, d_bytes = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;  // t_admit: taken into a slot's batch
  int status = -1;  // -1 pending/running; >= 0 final status
};

// Statistics and error state shared by the server and its tests.
struct Core {
  Pool pool;
  std::string err;  // message of the last failed slot
  std::vector<double> lat;
  int64_t ia_done = 0;
  uint64_t pairs_done = 0;
  double t_first = 0, t_last = 0;
};

// Completes one request: its final status, its pool memory released, and its
// latency counted when it succeeded.
inline void finish(Core& sv, Req* r, int status, double t) {
  r->status = status;
  r->t_done = t;
  if (r->d_off != SIZE_MAX) sv.pool.release(r->d_off, r->d_bytes);
  r->d_off = SIZE_MAX;
  if (status == MQPIV_OK) {
#ifdef MQPIV_SERVER_TRACE  // per-request latency lines on stderr (a measurement build)
    fprintf(stderr, "mqpiv-server-trace id=%llu ndims=%d n_ia=%lld submit=%.6f admit_ms=%.3f latency_ms=%.3f\n",
            (unsigned long long)r->id, (int)r->req.params.ndims, (long long)r->n_ia, r->t_submit,
            (r->t_admit - r->t_submit) * 1e3, (t - r->t_submit) * 1e3);
#endif
    sv.lat.push_back(t - r->t_submit);
    sv.ia_done += r->n_ia;
    sv.pairs_done += 1;
    if (sv.t_first == 0) sv.t_first = r->t_submit;
    sv.t_last = t;
  }
}

// What polling a busy slot's completion event found.
enum class Poll { kRunning, kDone, kFailed };

// Retires every busy slot whose batch has ended.  poll(slot, msg) reports the
// slot's state; kFailed stands for any query result other than "done" or "not
// ready" (a sticky device error after an asynchronous fault, an invalid
// event), and msg then says what it was.  The slot's requests finish with
// MQPIV_OK, or MQPIV_ERR_DEVICE on failure (recorded in err); either way their
// pool memory is released and the slot becomes idle, so that waiters and
// shutdown never block on a slot that cannot complete.  Returns whether any
// slot was retired.
template <class Slots, class PollFn>
bool retire(Core& sv, Slots& slots, double t, PollFn poll) {
  bool any = false;
  for (auto& sl : slots) {
    if (!sl.busy) continue;
    std::string msg;
    const Poll p = poll(sl, msg);
    if (p == Poll::kRunning) continue;
    const int status = p == Poll::kDone ? MQPIV_OK : MQPIV_ERR_DEVICE;
    if (status != MQPIV_OK) sv.err = msg.empty() ? std::string("device error in a running batch") : msg;
    for (Req* r : sl.reqs) finish(sv, r, status, t);
    sl.reqs.clear();
    sl.busy = false;
    any = true;
  }
  return any;
}

}  // namespace mqpiv_srv
