// Re-solve server (mccfr_server_* in mccfr.h): continuous batching of subgame
// re-solves on one context.  Requests wait in arrival order until their
// tables fit in the server's HBM pool (first-fit over one allocation, so
// admission never calls hipMalloc or hipFree) and a running place is free;
// then each is a tenant, and every step issues one batched launch holding the
// next launch of every running request.  At most two launches are in flight,
// so a request that arrives while others run joins within one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "runtime.hpp"

using namespace mccfr;

namespace {

constexpr size_t kMaxInFlight = 2;
constexpr size_t kPoolAlign = 256;

size_t align_up(size_t x) { return (x + kPoolAlign - 1) / kPoolAlign * kPoolAlign; }

// First-fit allocator over [0, size) with coalescing of free neighbours.
class Pool {
 public:
  void reset(size_t size) {
    free_.clear();
    if (size) free_[0] = size;
  }
  bool alloc(size_t n, size_t* off) {
    n = align_up(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const size_t rest = it->second - n, o = it->first + n;
      free_.erase(it);
      if (rest) free_[o] = rest;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = align_up(n);
    auto it = free_.emplace(off, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<size_t, size_t> free_;  // offset -> bytes
};

struct Request {
  mccfr_game_cfg cfg{};
  mccfr_root_state root{};
  int64_t iters = 0;
  uint32_t batch = 0;
  int state = 0;  // 0 waiting, 1 running, 2 done, 3 released, 4 failed
  bool failed = false;  // its launch no longer fits the context's launch scratch
  uint32_t tenant = 0;
  size_t off = 0;
  uint64_t bytes = 0;
  int64_t total = 0, issued = 0, finished = 0;
  uint64_t traversals = 0;
};

struct InFlight {
  hipEvent_t ev;
  std::vector<uint64_t> reqs;
};

}  // namespace

struct mccfr_server {
  mccfr_ctx* ctx = nullptr;
  void* pool_mem = nullptr;
  uint64_t budget = 0;
  uint32_t max_running = 0;
  Pool pool;
  std::map<uint64_t, Request> reqs;
  std::deque<uint64_t> waiting;
  std::vector<uint64_t> running;  // admitted requests with launches left to issue
  std::deque<InFlight> inflight;
  std::vector<hipEvent_t> spare;
  uint64_t next_id = 1;
};

namespace {

#define SRVX(srv, call)                                                                       \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) return fail((srv)->ctx, MCCFR_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Launch k of a request: round r = k / seats + 1 gives seat k % seats
// min(batch, iters - (r - 1) * batch) traversals.
HostSeg launch_of(const mccfr_ctx* c, const Request& q) {
  const Tenant& t = c->tenants[q.tenant];
  const int64_t k = q.issued;
  const uint32_t j = (uint32_t)(k % t.hg.n_upd);
  const int64_t r = k / t.hg.n_upd + 1;
  HostSeg s{};
  s.tenant = q.tenant;
  s.n = (uint32_t)std::min<int64_t>(q.batch, q.iters - (r - 1) * (int64_t)q.batch);
  s.update_player = t.hg.upd_seats[j];
  s.t_weight = t.cfg.use_linear_cfr ? r : 1;
  s.first_tid = t.traversals;
  s.iteration = r;
  s.upd_idx = j;
  return s;
}

int retire(mccfr_server* s, bool wait) {
  if (wait && !s->inflight.empty()) SRVX(s, hipEventSynchronize(s->inflight.front().ev));
  while (!s->inflight.empty()) {
    const hipError_t q = hipEventQuery(s->inflight.front().ev);
    if (q == hipErrorNotReady) break;
    if (q != hipSuccess) return fail(s->ctx, MCCFR_E_HIP, std::string("server launch: ") + hipGetErrorString(q));
    for (uint64_t id : s->inflight.front().reqs) {
      Request& r = s->reqs[id];
      if (++r.finished == r.total) r.state = r.failed ? 4 : 2;
    }
    s->spare.push_back(s->inflight.front().ev);
    s->inflight.pop_front();
  }
  return MCCFR_OK;
}

int admit(mccfr_server* s) {
  while (!s->waiting.empty() && s->running.size() < s->max_running) {
    Request& r = s->reqs[s->waiting.front()];
    size_t off = 0;
    if (!s->pool.alloc(r.bytes, &off)) break;  // first come, first served
    const int rc = tenant_create(s->ctx, &r.cfg, &r.root, static_cast<char*>(s->pool_mem) + off, &r.tenant);
    if (rc) {
      s->pool.release(off, r.bytes);
      return rc;
    }
    r.off = off;
    r.state = 1;
    r.total = (r.iters + r.batch - 1) / r.batch * (int64_t)s->ctx->tenants[r.tenant].hg.n_upd;
    s->running.push_back(s->waiting.front());
    s->waiting.pop_front();
  }
  return MCCFR_OK;
}

int issue(mccfr_server* s) {
  if (s->inflight.size() >= kMaxInFlight || s->running.empty()) return MCCFR_OK;
  mccfr_ctx* c = s->ctx;
  std::vector<HostSeg> segs;
  std::vector<uint64_t> ids;
  uint64_t lanes = 0;
  int row = -1;
  std::string failed;
  for (uint64_t id : s->running) {
    Request& r = s->reqs[id];
    if (r.issued >= r.total) continue;
    const Tenant& t = c->tenants[r.tenant];
    const HostSeg hs = launch_of(c, r);
    const uint64_t padded = (hs.n + kSegAlign - 1) / kSegAlign * kSegAlign;
    if (padded > c->sc.cap_lanes) {
      // the context's launch scratch shrank below this request's launch
      // (mccfr_set_max_batch): it can never be issued, so it fails rather
      // than wait for ever; launches already in flight finish first
      r.failed = true;
      r.total = r.issued;
      if (r.finished == r.total) r.state = 4;
      failed += (failed.empty() ? "" : ", ") + std::to_string(id);
      continue;
    }
    if (row < 0) row = t.row;
    if (t.row != row) continue;  // other row widths go in a later launch
    if (lanes + padded > c->sc.cap_lanes) continue;
    lanes += padded;
    segs.push_back(hs);
    ids.push_back(id);
  }
  const auto report = [&](int rc) {
    if (failed.empty()) return rc;
    s->running.erase(std::remove_if(s->running.begin(), s->running.end(),
                                    [&](uint64_t id) { return s->reqs[id].failed; }),
                     s->running.end());
    return rc ? rc
              : fail(c, MCCFR_E_STATE,
                     "request(s) " + failed + " failed: launch larger than the context's max_batch");
  };
  if (segs.empty()) return report(MCCFR_OK);
  hipEvent_t ev = nullptr;
  if (!s->spare.empty()) {
    ev = s->spare.back();
    s->spare.pop_back();
  } else {
    SRVX(s, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  const int rc = issue_launch(c, segs.data(), (uint32_t)segs.size(), nullptr, ev);
  if (rc) {
    s->spare.push_back(ev);
    return report(rc);
  }
  for (size_t i = 0; i < ids.size(); ++i) {
    Request& r = s->reqs[ids[i]];
    ++r.issued;
    r.traversals += segs[i].n;
  }
  s->inflight.push_back(InFlight{ev, ids});
  s->running.erase(std::remove_if(s->running.begin(), s->running.end(),
                                  [&](uint64_t id) { const Request& r = s->reqs[id]; return r.issued >= r.total; }),
                   s->running.end());
  return report(MCCFR_OK);
}

Request* find(mccfr_server* s, uint64_t id) {
  auto it = s->reqs.find(id);
  return it == s->reqs.end() ? nullptr : &it->second;
}

}  // namespace

extern "C" {

int mccfr_server_create(mccfr_ctx* c, uint64_t hbm_budget, uint32_t max_running, mccfr_server** out) {
  if (!c || !out) return fail(c, MCCFR_E_ARG, "null argument");
  *out = nullptr;
  if (hbm_budget < kPoolAlign) return fail(c, MCCFR_E_ARG, "hbm_budget too small");
  if (max_running == 0 || max_running > kMaxSegs) return fail(c, MCCFR_E_ARG, "max_running must be 1..4096");
  auto* s = new (std::nothrow) mccfr_server();
  if (!s) return fail(c, MCCFR_E_NOMEM, "host allocation failed");
  s->ctx = c;
  s->budget = align_up(hbm_budget);
  s->max_running = max_running;
  const hipError_t e = hipMalloc(&s->pool_mem, s->budget);
  if (e != hipSuccess) {
    delete s;
    return fail(c, MCCFR_E_NOMEM, std::string("server pool: ") + hipGetErrorString(e));
  }
  s->pool.reset(s->budget);
  *out = s;
  return MCCFR_OK;
}

void mccfr_server_destroy(mccfr_server* s) {
  if (!s) return;
  for (auto& f : s->inflight) {
    (void)hipEventSynchronize(f.ev);
    (void)hipEventDestroy(f.ev);
  }
  for (hipEvent_t e : s->spare) (void)hipEventDestroy(e);
  for (auto& kv : s->reqs)
    if (kv.second.state == 1 || kv.second.state == 2 || kv.second.state == 4) (void)tenant_destroy(s->ctx, kv.second.tenant);
  (void)hipStreamSynchronize(s->ctx->stream);
  (void)hipFree(s->pool_mem);
  delete s;
}

int mccfr_server_submit(mccfr_server* s, const mccfr_game_cfg* cfg, const mccfr_root_state* root, int64_t iterations,
                        uint32_t batch, uint64_t* request) {
  if (!s || !cfg || !request) return fail(s ? s->ctx : nullptr, MCCFR_E_ARG, "null argument");
  DevGame g;
  const char* why = nullptr;
  if (!make_dev_game(*cfg, root, g, &why)) return fail(s->ctx, MCCFR_E_ARG, why);
  if (cfg->sampling != MCCFR_SAMPLE_OUTCOME)
    return fail(s->ctx, MCCFR_E_ARG, "the server runs outcome-sampling requests (MCCFR_SAMPLE_OUTCOME)");
  if (iterations < 1) return fail(s->ctx, MCCFR_E_ARG, "iterations must be >= 1");
  if (batch < 1 || batch > s->ctx->max_batch) return fail(s->ctx, MCCFR_E_ARG, "batch must be 1..max_batch");
  Request r;
  r.cfg = *cfg;
  if (root) r.root = *root;
  r.iters = iterations;
  r.batch = batch;
  r.bytes = tenant_bytes(*cfg);
  if (align_up(r.bytes) > s->budget) return fail(s->ctx, MCCFR_E_ARG, "request tables exceed the server's HBM budget");
  const uint64_t id = s->next_id++;
  s->reqs[id] = r;
  s->waiting.push_back(id);
  *request = id;
  return MCCFR_OK;
}

int mccfr_server_step(mccfr_server* s, int wait, uint32_t* running, uint32_t* waiting) {
  if (!s) return fail(nullptr, MCCFR_E_ARG, "null server");
  int rc = retire(s, wait != 0);
  if (!rc) rc = admit(s);
  if (!rc) rc = issue(s);
  if (rc) return rc;
  if (running) {
    uint32_t n = 0;
    for (auto& kv : s->reqs) n += kv.second.state == 1;
    *running = n;
  }
  if (waiting) *waiting = (uint32_t)s->waiting.size();
  return MCCFR_OK;
}

int mccfr_server_poll(mccfr_server* s, uint64_t id, int* state, uint64_t* traversals) {
  if (!s) return fail(nullptr, MCCFR_E_ARG, "null server");
  const Request* r = find(s, id);
  if (!r) return fail(s->ctx, MCCFR_E_ARG, "no such request");
  if (state) *state = r->state;
  if (traversals) *traversals = r->traversals;
  return MCCFR_OK;
}

int mccfr_server_tenant(mccfr_server* s, uint64_t id, uint32_t* tenant) {
  if (!s || !tenant) return fail(s ? s->ctx : nullptr, MCCFR_E_ARG, "null argument");
  const Request* r = find(s, id);
  if (!r || r->state != 2) return fail(s->ctx, MCCFR_E_STATE, "request is not done");
  *tenant = r->tenant;
  return MCCFR_OK;
}

int mccfr_server_result(mccfr_server* s, uint64_t id, float* strategy, size_t elems, int layout) {
  if (!s) return fail(nullptr, MCCFR_E_ARG, "null server");
  const Request* r = find(s, id);
  if (!r || r->state != 2) return fail(s->ctx, MCCFR_E_STATE, "request is not done");
  return tenant_normalized(s->ctx, r->tenant, strategy, elems, layout);
}

int mccfr_server_release(mccfr_server* s, uint64_t id) {
  if (!s) return fail(nullptr, MCCFR_E_ARG, "null server");
  Request* r = find(s, id);
  if (!r || (r->state != 2 && r->state != 4)) return fail(s->ctx, MCCFR_E_STATE, "request is not done");
  const int rc = tenant_destroy(s->ctx, r->tenant);
  if (rc) return rc;
  s->pool.release(r->off, r->bytes);
  r->state = 3;
  return MCCFR_OK;
}

}  // extern "C"
