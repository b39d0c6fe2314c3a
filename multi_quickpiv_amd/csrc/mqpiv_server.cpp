// mqpiv_server.cpp -- continuous-batching request loop over libmqpiv.
//
// Clients submit frame pairs (host buffers, or device buffers already resident
// in HBM) of any size, dimensionality and parameters; a worker thread admits
// them into free execution slots as device memory allows (admission by
// footprint against a fixed HBM pool that also holds every pair's integral
// arrays), stages host inputs in, runs each admitted set as one batched call
// (one kernel launch per algorithm / precision / transform-length group of its
// 2-D pairs) and copies host outputs back.  Several slots, each with its own
// context and stream, are in flight at once, so copies and kernels of different
// batches overlap and requests arriving while others run are admitted as soon
// as a slot and memory free up.  Every IA is computed independently of the
// batch it lands in, so a request's results are the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mqpiv.h"
#include "mqpiv_batch.h"
#include "mqpiv_server_core.h"

namespace {

using mqpiv_srv::al256;
using mqpiv_srv::Req;

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Slot {
  mqpiv_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<Req*> reqs;
  bool busy = false;
  // completion events of the batches this one must follow on the device (taken
  // at admission from slots still running then)
  std::vector<hipEvent_t> after;
};

}  // namespace

struct mqpiv_server : mqpiv_srv::Core {
  int device = 0;
  char* d_pool = nullptr;
  int max_batch = 64;
  std::vector<Slot> slots;
  std::deque<Req*> pending;
  std::unordered_map<uint64_t, Req*> reqs;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread worker;
  bool stop = false;
  uint64_t next_id = 1;
  size_t next_slot = 0;
};

// Device bytes a request holds while admitted: for host requests its staged
// inputs and outputs, and for every request its integral arrays.
static void size_request(Req& r) {
  const mqpiv_request& q = r.req;
  const int nd = q.params.ndims;
  size_t io = 0;
  if (!r.on_device) {
    io += al256(r.npx * r.esz) * (q.params.corr_alg == MQPIV_MASK_NSQECC ? 3 : 2);
    io += al256((size_t)r.n_ia * nd * r.osz);
    if (q.SN) io += al256((size_t)r.n_ia * r.osz);
    if (q.peaks) io += al256((size_t)r.n_ia * nd * sizeof(int32_t));
  }
  r.io_bytes = io;
  r.ws_bytes = al256(request_ws_bytes(q));
  r.d_bytes = io + r.ws_bytes;
}

// Issues copies in (host requests), the batched computation and copies out for
// the requests of one slot; returns a status for the whole slot.
static int launch_slot(mqpiv_server* sv, Slot& sl) {
  const size_t n = sl.reqs.size();
  std::vector<mqpiv_request> dev(n);
  std::vector<void*> ws(n, nullptr);
  for (size_t i = 0; i < n; ++i) {
    Req& r = *sl.reqs[i];
    char* p = sv->d_pool + r.d_off;
    mqpiv_request& d = dev[i];
    d = r.req;
    d.skip_mask = nullptr;  // host-computed skip masks are not taken by the server
    if (!r.on_device) {
      const size_t ib = r.npx * r.esz;
      if (hipMemcpyAsync(p, r.req.img1, ib, hipMemcpyHostToDevice, sl.stream) != hipSuccess) return MQPIV_ERR_DEVICE;
      d.img1 = p;
      p += al256(ib);
      if (hipMemcpyAsync(p, r.req.img2, ib, hipMemcpyHostToDevice, sl.stream) != hipSuccess) return MQPIV_ERR_DEVICE;
      d.img2 = p;
      p += al256(ib);
      if (r.req.params.corr_alg == MQPIV_MASK_NSQECC) {
        if (hipMemcpyAsync(p, r.req.mask, ib, hipMemcpyHostToDevice, sl.stream) != hipSuccess)
          return MQPIV_ERR_DEVICE;
        d.mask = p;
        p += al256(ib);
      }
      d.VF = p;
      p += al256((size_t)r.n_ia * r.req.params.ndims * r.osz);
      if (r.req.SN) {
        d.SN = p;
        p += al256((size_t)r.n_ia * r.osz);
      }
      if (r.req.peaks) {
        d.peaks = (int32_t*)p;
        p += al256((size_t)r.n_ia * r.req.params.ndims * sizeof(int32_t));
      }
    }
    if (r.ws_bytes) ws[i] = sv->d_pool + r.d_off + r.io_bytes;
  }
  // the inputs are staged; the kernels start once the batches ahead are done
  for (hipEvent_t e : sl.after)
    if (hipStreamWaitEvent(sl.stream, e, 0) != hipSuccess) return MQPIV_ERR_DEVICE;
  int rc = batch_device_ws(sl.ctx, dev.data(), (int32_t)n, sl.stream, ws.data());
  if (rc) return rc;
  for (size_t i = 0; i < n; ++i) {
    Req& r = *sl.reqs[i];
    if (r.on_device) continue;
    const mqpiv_request& d = dev[i];
    if (hipMemcpyAsync(r.req.VF, d.VF, (size_t)r.n_ia * r.req.params.ndims * r.osz, hipMemcpyDeviceToHost,
                       sl.stream) != hipSuccess)
      return MQPIV_ERR_DEVICE;
    if (r.req.SN && hipMemcpyAsync(r.req.SN, d.SN, (size_t)r.n_ia * r.osz, hipMemcpyDeviceToHost, sl.stream) !=
                        hipSuccess)
      return MQPIV_ERR_DEVICE;
    if (r.req.peaks &&
        hipMemcpyAsync(r.req.peaks, d.peaks, (size_t)r.n_ia * r.req.params.ndims * sizeof(int32_t),
                       hipMemcpyDeviceToHost, sl.stream) != hipSuccess)
      return MQPIV_ERR_DEVICE;
  }
  if (hipEventRecord(sl.done, sl.stream) != hipSuccess) return MQPIV_ERR_DEVICE;
  return MQPIV_OK;
}

static void finish(mqpiv_server* sv, Req* r, int status, double t) { mqpiv_srv::finish(*sv, r, status, t); }

// A request that fills the device by itself: a 3-D volume, or a 2-D pair of more
// than kBigIA IAs (a 2048 x 2048 pair at interSize 16, 16,129 IAs, is not).
static constexpr int64_t kBigIA = 32768;
static bool is_big(const Req& r) { return r.req.params.ndims == 3 || r.n_ia > kBigIA; }

static void worker_loop(mqpiv_server* sv) {
  hipSetDevice(sv->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  while (true) {
    bool progress = false;
    // retire finished slots; a failed query ends the slot's requests with a
    // device error instead of leaving the slot busy for ever
    if (mqpiv_srv::retire(*sv, sv->slots, now_s(), [](Slot& sl, std::string& msg) {
          const hipError_t e = hipEventQuery(sl.done);
          if (e == hipErrorNotReady) return mqpiv_srv::Poll::kRunning;
          if (e == hipSuccess) return mqpiv_srv::Poll::kDone;
          msg = std::string("device error in a running batch: ") + hipGetErrorString(e);
          return mqpiv_srv::Poll::kFailed;
        })) {
      progress = true;
      sv->cv_done.notify_all();
    }
    if (sv->stop && sv->pending.empty()) {
      bool any = false;
      for (auto& sl : sv->slots) any |= sl.busy;
      if (!any) break;
    }
    // admit pending requests into free slots (round-robin) in arrival order,
    // while memory lasts.  The device runs them in that order as well: a large
    // request (a 3-D volume, or a 2-D pair of more than kBigIA IAs, which fills
    // the device by itself) runs alone -- work beside it only takes CUs from it,
    // and the requests of a batch complete together, so anything sharing its
    // launch or the device would delay it.  Small requests form batches of the
    // consecutive small requests of one dimensionality from the front (their
    // launches would leave most CUs idle alone), which may run beside other
    // small batches.  The order is kept on the device, not by waiting here: a
    // batch's stream waits for the completion events of the slots it must
    // follow (every running slot for a large request, the running large
    // request for a small batch), so it starts the moment they end.
    const size_t ns = sv->slots.size();
    for (size_t k = 0; k < ns; ++k) {
      Slot& sl = sv->slots[(sv->next_slot + k) % ns];
      if (sl.busy || sv->pending.empty()) continue;
      const Req& front = *sv->pending.front();
      const bool big = is_big(front);
      const int front_nd = front.req.params.ndims;
      for (auto it = sv->pending.begin(); it != sv->pending.end() && (int)sl.reqs.size() < sv->max_batch;) {
        Req* r = *it;
        if (!big && (is_big(*r) || r->req.params.ndims != front_nd)) break;
        const size_t off = sv->pool.alloc(r->d_bytes);
        if (off == SIZE_MAX) break;  // wait for memory to free up
        r->d_off = off;
        it = sv->pending.erase(it);
        sl.reqs.push_back(r);
        if (big) break;  // a large request runs alone
      }
      sl.after.clear();
      for (auto& s2 : sv->slots)
        if (s2.busy && (big || (s2.reqs.size() == 1 && is_big(*s2.reqs.front())))) sl.after.push_back(s2.done);
      if (sl.reqs.empty()) continue;
      const double t_admit = now_s();
      for (Req* r : sl.reqs) r->t_admit = t_admit;
      // copies and launches run without the lock so that submitters are not held up
      lk.unlock();
      const int rc = launch_slot(sv, sl);
      if (rc != MQPIV_OK) hipStreamSynchronize(sl.stream);
      lk.lock();
      if (rc != MQPIV_OK) {
        const double t = now_s();
        sv->err = mqpiv_last_error(sl.ctx);
        if (sv->err.empty()) sv->err = "HIP runtime error while staging a batch";
        for (Req* r : sl.reqs) finish(sv, r, rc, t);
        sl.reqs.clear();
        sv->cv_done.notify_all();
      } else {
        sl.busy = true;
      }
      sv->next_slot = (sv->next_slot + k + 1) % ns;
      progress = true;
      break;  // re-check completions before filling the next slot
    }
    if (progress) continue;
    bool any_busy = false;
    for (auto& s2 : sv->slots) any_busy |= s2.busy;
    if (any_busy) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lk.lock();
    } else {
      sv->cv_work.wait(lk);
    }
  }
}

extern "C" int mqpiv_server_create(int device, uint64_t hbm_budget_bytes, int32_t n_slots, int32_t max_batch,
                                   mqpiv_server** out) {
  if (!out) return MQPIV_ERR_NULL;
  *out = nullptr;
  if (n_slots < 1 || n_slots > 16 || max_batch < 1 || hbm_budget_bytes < (1u << 20)) return MQPIV_ERR_BAD_PARAMS;
  if (hipSetDevice(device) != hipSuccess) return MQPIV_ERR_DEVICE;
  mqpiv_server* sv = new mqpiv_server();
  sv->device = device;
  sv->max_batch = max_batch;
  if (hipMalloc((void**)&sv->d_pool, hbm_budget_bytes) != hipSuccess) {
    delete sv;
    return MQPIV_ERR_OOM;
  }
  sv->pool.init(hbm_budget_bytes);
  sv->slots.resize(n_slots);
  for (auto& sl : sv->slots) {
    int rc = mqpiv_ctx_create(device, 0, &sl.ctx);
    if (rc == MQPIV_OK && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess)
      rc = MQPIV_ERR_DEVICE;
    if (rc == MQPIV_OK && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
      rc = MQPIV_ERR_DEVICE;
    if (rc != MQPIV_OK) {
      for (auto& s2 : sv->slots) {
        if (s2.ctx) mqpiv_ctx_destroy(s2.ctx);
        if (s2.stream) hipStreamDestroy(s2.stream);
        if (s2.done) hipEventDestroy(s2.done);
      }
      hipFree(sv->d_pool);
      delete sv;
      return rc;
    }
  }
  sv->worker = std::thread(worker_loop, sv);
  *out = sv;
  return MQPIV_OK;
}

static int submit(mqpiv_server* sv, const mqpiv_request* req, bool on_device, uint64_t* ticket) {
  if (!sv || !req || !ticket) return MQPIV_ERR_NULL;
  int64_t vf[3], n_ia = 0;
  int rc = mqpiv_vf_size(&req->params, req->dims, vf, &n_ia);
  if (rc) return rc;
  if (!req->img1 || !req->img2 || (n_ia > 0 && !req->VF)) return MQPIV_ERR_NULL;
  if (req->params.corr_alg == MQPIV_MASK_NSQECC && !req->mask) return MQPIV_ERR_NULL;
  if (req->params.compute_sn && !req->SN) return MQPIV_ERR_NULL;
  if (req->skip_mask) return MQPIV_ERR_UNSUPPORTED;
  if ((req->in_dtype != MQPIV_F32 && req->in_dtype != MQPIV_F64) ||
      (req->out_dtype != MQPIV_F32 && req->out_dtype != MQPIV_F64))
    return MQPIV_ERR_BAD_PARAMS;
  Req* r = new Req();
  r->req = *req;
  r->on_device = on_device;
  r->n_ia = n_ia;
  r->npx = (size_t)req->dims[0] * (size_t)req->dims[1] * (size_t)(req->params.ndims == 3 ? req->dims[2] : 1);
  r->esz = req->in_dtype == MQPIV_F32 ? 4 : 8;
  r->osz = req->out_dtype == MQPIV_F32 ? 4 : 8;
  size_request(*r);
  r->t_submit = now_s();
  std::lock_guard<std::mutex> lk(sv->mu);
  if (r->d_bytes > sv->pool.capacity()) {
    delete r;
    return MQPIV_ERR_OOM;  // larger than the whole budget: rejected at admission
  }
  r->id = sv->next_id++;
  *ticket = r->id;
  sv->reqs[r->id] = r;
  if (n_ia == 0) {
    finish(sv, r, MQPIV_OK, now_s());
  } else {
    sv->pending.push_back(r);
    sv->cv_work.notify_one();
  }
  return MQPIV_OK;
}

extern "C" int mqpiv_server_submit(mqpiv_server* sv, const mqpiv_request* req, uint64_t* ticket) {
  return submit(sv, req, false, ticket);
}

extern "C" int mqpiv_server_submit_device(mqpiv_server* sv, const mqpiv_request* req, uint64_t* ticket) {
  return submit(sv, req, true, ticket);
}

extern "C" int mqpiv_server_wait(mqpiv_server* sv, uint64_t ticket, double timeout_s, int32_t* status) {
  if (!sv || !status) return MQPIV_ERR_NULL;
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return MQPIV_ERR_BAD_PARAMS;
  Req* r = it->second;
  auto pred = [&] { return r->status >= 0; };
  if (timeout_s < 0) {
    sv->cv_done.wait(lk, pred);
  } else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
    *status = -1;  // still running
    return MQPIV_OK;
  }
  *status = r->status;
  sv->reqs.erase(it);
  delete r;
  return MQPIV_OK;
}

extern "C" int mqpiv_server_stats_get(mqpiv_server* sv, mqpiv_server_stats* st) {
  if (!sv || !st) return MQPIV_ERR_NULL;
  std::lock_guard<std::mutex> lk(sv->mu);
  memset(st, 0, sizeof(*st));
  st->completed = sv->pairs_done;
  st->ia_done = sv->ia_done;
  st->pending = sv->pending.size();
  if (!sv->lat.empty()) {
    std::vector<double> v = sv->lat;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
      const double pos = q * (v.size() - 1);
      const size_t i = (size_t)pos;
      const double f = pos - i;
      return (i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i]) * 1e3;
    };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = v.back() * 1e3;
    const double span = sv->t_last - sv->t_first;
    if (span > 0) {
      st->pairs_per_s = sv->pairs_done / span;
      st->ia_per_s = sv->ia_done / span;
    }
  }
  return MQPIV_OK;
}

extern "C" const char* mqpiv_server_last_error(mqpiv_server* sv) {
  if (!sv) return "null server";
  std::lock_guard<std::mutex> lk(sv->mu);
  thread_local std::string copy;
  copy = sv->err;
  return copy.c_str();
}

extern "C" int mqpiv_server_reset_stats(mqpiv_server* sv) {
  if (!sv) return MQPIV_ERR_NULL;
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->lat.clear();
  sv->ia_done = 0;
  sv->pairs_done = 0;
  sv->t_first = sv->t_last = 0;
  return MQPIV_OK;
}

extern "C" int mqpiv_server_destroy(mqpiv_server* sv) {
  if (!sv) return MQPIV_ERR_NULL;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    sv->cv_work.notify_all();
  }
  if (sv->worker.joinable()) sv->worker.join();
  for (auto& sl : sv->slots) {
    if (sl.stream) hipStreamSynchronize(sl.stream);
    if (sl.ctx) mqpiv_ctx_destroy(sl.ctx);
    if (sl.stream) hipStreamDestroy(sl.stream);
    if (sl.done) hipEventDestroy(sl.done);
  }
  for (auto& kv : sv->reqs) delete kv.second;
  hipFree(sv->d_pool);
  delete sv;
  return MQPIV_OK;
}
