// CPU test of the request server's slot retirement (mqpiv_server_core.h):
// polling results are injected per slot, and the outcome is printed as
// "name value" lines for tests/test_server_core.py.
#include <cstdio>
#include <string>
#include <vector>

#include "../multi_quickpiv_amd/csrc/mqpiv_server_core.h"

using namespace mqpiv_srv;

struct FakeSlot {
  bool busy = false;
  std::vector<Req*> reqs;
  Poll next = Poll::kRunning;  // what the next poll reports
};

int main() {
  Core sv;
  sv.pool.init(1 << 20);
  std::vector<Req> rq(4);
  std::vector<FakeSlot> slots(3);
  for (int i = 0; i < 4; ++i) {
    rq[i].id = i + 1;
    rq[i].n_ia = 100 * (i + 1);
    rq[i].d_bytes = 4096 * (i + 1);
    rq[i].d_off = sv.pool.alloc(rq[i].d_bytes);
    rq[i].t_submit = 1.0;
  }
  // slot 0: requests 0 and 1, its event failed; slot 1: request 2, done;
  // slot 2: request 3, still running
  slots[0].busy = true;
  slots[0].reqs = {&rq[0], &rq[1]};
  slots[0].next = Poll::kFailed;
  slots[1].busy = true;
  slots[1].reqs = {&rq[2]};
  slots[1].next = Poll::kDone;
  slots[2].busy = true;
  slots[2].reqs = {&rq[3]};
  slots[2].next = Poll::kRunning;
  std::printf("used_before %zu\n", sv.pool.used());
  int polls = 0;
  auto poll = [&](FakeSlot& s, std::string& msg) {
    ++polls;
    if (s.next == Poll::kFailed) msg = "injected: illegal address";
    return s.next;
  };
  const bool any = retire(sv, slots, 2.5, poll);
  std::printf("retired_any %d\n", any ? 1 : 0);
  std::printf("polls %d\n", polls);
  for (int i = 0; i < 4; ++i) std::printf("status%d %d\n", i, rq[i].status);
  for (int i = 0; i < 3; ++i) std::printf("busy%d %d\n", i, slots[i].busy ? 1 : 0);
  std::printf("used_after %zu\n", sv.pool.used());
  std::printf("pairs_done %llu\n", (unsigned long long)sv.pairs_done);
  std::printf("ia_done %lld\n", (long long)sv.ia_done);
  std::printf("err_set %d\n", sv.err == "injected: illegal address" ? 1 : 0);
  // the running slot then fails too: every request has ended and nothing is busy
  slots[2].next = Poll::kFailed;
  retire(sv, slots, 3.0, poll);
  std::printf("status3_later %d\n", rq[3].status);
  std::printf("busy2_later %d\n", slots[2].busy ? 1 : 0);
  std::printf("used_final %zu\n", sv.pool.used());
  // the pool coalesced back into one free block: the whole capacity allocates again
  std::printf("realloc_whole %d\n", sv.pool.alloc(1 << 20) == 0 ? 1 : 0);
  // an idle slot is never polled
  polls = 0;
  retire(sv, slots, 4.0, poll);
  std::printf("polls_idle %d\n", polls);
  return 0;
}
