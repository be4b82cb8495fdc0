// Host runtime internals shared by runtime.hip (contexts, tenants, launches)
// and serve.hip (the re-solve server).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mccfr/mccfr.h"
#include "device_common.hpp"
#include "kernels.hpp"

namespace mccfr {

constexpr uint32_t kMaxTenants = 4096;
constexpr uint32_t kMaxSegs = 4096;
constexpr uint32_t kStageRing = 8;       // pinned staging slots for segment tables
constexpr uint32_t kTimingRing = 1024;   // timed launches remembered

struct Tenant {
  bool used = false;
  mccfr_game_cfg cfg{};
  DevGame hg{};           // host copy of the device descriptor (with table pointers)
  uint64_t slots = 0;
  int row = 8;
  void* owned = nullptr;  // table allocation owned by the tenant (null: caller-provided memory)
  bool sigma_dirty = false;
  bool clamp_pending = false;  // tables uploaded with CFR+ on: clamp all after the next launch
  uint64_t traversals = 0;     // traversals issued = next traversal id of the schedule
  uint64_t launches = 0;
  int64_t next_iter = 1;
  uint32_t upd_idx = 0;
  size_t elems() const { return (size_t)slots * row; }
};

// One segment as the host issues it.
struct HostSeg {
  uint32_t tenant;
  uint32_t n;
  int32_t update_player;
  int64_t t_weight;
  uint64_t first_tid;
  int64_t iteration;
  uint32_t upd_idx;
};

}  // namespace mccfr

struct mccfr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;

  // shared read-only inputs
  int32_t* hr = nullptr;
  uint8_t* lut = nullptr;
  mccfr::ArithTables arith{};
  mccfr::BucketTables buckets{};
  mccfr::BucketTables eval_bk{};  // the strategy harness's clamped buckets, built on first use
  bool eval_bk_ready = false;

  // tenants
  std::vector<mccfr::Tenant> tenants;
  mccfr::DevGame* d_tenants = nullptr;

  // launch scratch
  uint32_t max_batch = 0;
  mccfr::DevScratch sc{};
  mccfr::DevSeg* d_segs = nullptr;
  mccfr_trace* dtrace = nullptr;
  uint32_t* es_stack = nullptr;  // external-sampling walk frames, grown to the largest such launch
  size_t es_stack_bytes = 0;
  float* es_rd = nullptr;  // external-sampling regret-delta rows, one per sort-input record
  size_t es_rd_bytes = 0;
  char* es_gen = nullptr;  // external sampling by generations: tasks, nodes, trace sums
  size_t es_gen_size = 0;
  uint32_t es_gen_cap = 0;  // the scratch capacity its carving in sc.gen was made for
  mccfr::EsGenBufs es_gen_bufs{};
  mccfr::DevSeg* h_stage = nullptr;  // pinned [kStageRing][kMaxSegs]
  hipEvent_t stage_ev[mccfr::kStageRing] = {};
  uint32_t stage_next = 0;
  mccfr::WalkShape shape{};

  // captured training graph and what it was captured for
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> graph_key;

  uint64_t traversals = 0, launches = 0;

  bool timing = false;
  std::vector<hipEvent_t> ev;  // 4 per timed launch: start, after deal, after walk, after table update
  uint32_t ev_next = 0, ev_used = 0;

  std::string err;
};

namespace mccfr {

int fail(mccfr_ctx* c, int code, const std::string& msg);

// Returns from the calling entry point with the HIP error recorded.
#define HIPX(ctx, call)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ::mccfr::fail((ctx), e_ == hipErrorOutOfMemory ? MCCFR_E_NOMEM : MCCFR_E_HIP,  \
                           std::string(#call) + ": " + hipGetErrorString(e_));              \
  } while (0)
// Validates a game configuration (+ root) and fills the device descriptor
// (without table pointers).
bool make_dev_game(const mccfr_game_cfg& cfg, const mccfr_root_state* root, DevGame& g, const char** why);
uint64_t tenant_bytes(const mccfr_game_cfg& cfg);
// Creates a tenant; `mem` (tenant_bytes of device memory) or null to allocate.
int tenant_create(mccfr_ctx* c, const mccfr_game_cfg* cfg, const mccfr_root_state* root, void* mem,
                  uint32_t* id);
int tenant_destroy(mccfr_ctx* c, uint32_t id);
// Issues one batched launch; `done` (may be null) is recorded after it.
// `scratch`: launch scratch other than the context's (null = the context's);
// such a launch takes no host trace.
int issue_launch(mccfr_ctx* c, const HostSeg* segs, uint32_t n, mccfr_trace* host_trace, hipEvent_t done,
                 const DevScratch* scratch = nullptr);
int tenant_normalized(mccfr_ctx* c, uint32_t id, float* out, size_t elems, int layout);

}  // namespace mccfr
