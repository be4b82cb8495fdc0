// Host runtime behind the C ABI (include/mccfr/mccfr.h): contexts, device
// buffers, table upload/download, launch sequencing and the hipGraph
// training loop.  Replaces GPUCFRTrainer's host side (cfr_gpu.cu:1064-1538).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mccfr/mccfr.h"
#include "device_common.hpp"
#include "handranks.hpp"
#include "host_tables.hpp"
#include "kernels.hpp"

using namespace mccfr;

struct mccfr_ctx {
  mccfr_game_cfg cfg{};
  DevGame hg{};
  DevGame* dg = nullptr;
  int row = 8;
  uint64_t slots = 0;
  uint32_t max_batch = 0;
  int device = 0;
  hipStream_t stream = nullptr;

  DevTables t{};
  LaunchState* st = nullptr;
  int32_t* hr = nullptr;
  uint8_t* lut = nullptr;
  ArithTables arith{};
  BucketTables buckets{};
  mccfr_trace* dtrace = nullptr;

  uint64_t traversals = 0, launches = 0;
  int64_t next_iter = 1;
  bool sigma_dirty = false;

  hipGraphExec_t gexec = nullptr;
  uint32_t graph_batch = 0;
  uint32_t walk_blocks = 0;   // persistent walk workgroups per launch (0: one lane per traversal)
  uint32_t scatter_wg = 1024; // scatter workgroup size (4 traversals per thread)

  bool timing = false;
  std::vector<hipEvent_t> ev;  // 3 events per timed launch: start, after traverse, after update
  uint32_t ev_next = 0, ev_used = 0;

  std::string err;
};

namespace {

constexpr uint32_t kTimingRing = 1024;
std::string g_err;
std::mutex g_err_mu;

int fail(mccfr_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = msg;
  return code;
}

#define HIPX(ctx, call)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail((ctx), e_ == hipErrorOutOfMemory ? MCCFR_E_NOMEM : MCCFR_E_HIP,           \
                  std::string(#call) + ": " + hipGetErrorString(e_));                       \
  } while (0)

const std::vector<int32_t>* generated_table() {
  static std::once_flag once;
  static std::vector<int32_t>* tbl = nullptr;
  std::call_once(once, []() {
    auto* v = new (std::nothrow) std::vector<int32_t>();
    if (!v) return;
    v->resize(kHandTableWords);
    if (generate_hand_table(v->data(), 0) != kHandStates) {
      delete v;
      return;
    }
    tbl = v;
  });
  return tbl;
}

bool table_ok(const int32_t* hr, size_t words) {  // hand_eval.cpp:59-86
  if (words != kHandTableWords) return false;
  const uint8_t royal[7] = {51, 50, 49, 48, 47, 0, 14};  // As Ks Qs Js Ts 2c 3d
  int p = hr[53 + card_rm(royal[0])];
  for (int i = 1; i < 6; ++i) {
    if (p < 0 || (size_t)p + 53 > words) return false;
    p = hr[p + card_rm(royal[i])];
  }
  if (p < 0 || (size_t)p + 53 > words) return false;
  return decode_packed((uint32_t)hr[p + card_rm(royal[6])] & 0xFFFFu) == 7462u;
}

// Per-launch node lists for `lanes` traversals (kMaxDepth entries per lane).
hipError_t alloc_node_lists(DevTables& t, uint32_t lanes) {
  hipError_t e = hipMalloc(&t.node_slot, (size_t)kMaxDepth * lanes * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&t.node_av, (size_t)kMaxDepth * lanes * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&t.node_count, (size_t)lanes);
  if (e == hipSuccess) e = hipMalloc(&t.node_u, (size_t)lanes * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&t.root_bkt, (size_t)kMaxPlayers * lanes * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&t.root_rank, (size_t)kMaxPlayers * lanes * sizeof(uint16_t));
  if (e == hipSuccess) e = hipMalloc(&t.root_dealer, (size_t)lanes);
  if (e == hipSuccess) e = hipMalloc(&t.upd_list, (size_t)kMaxDepth * lanes * sizeof(uint32_t));
  t.lanes = lanes;
  return e;
}

void free_node_lists(DevTables& t) {
  void* ptrs[] = {t.node_slot, t.node_av, t.node_count, t.node_u, t.root_bkt, t.root_rank, t.root_dealer,
                  t.upd_list};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  t.node_slot = nullptr;
  t.node_av = nullptr;
  t.node_count = nullptr;
  t.node_u = nullptr;
  t.root_bkt = nullptr;
  t.root_rank = nullptr;
  t.root_dealer = nullptr;
  t.upd_list = nullptr;
}

void free_all(mccfr_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  free_node_lists(c->t);
  void* ptrs[] = {c->dg, c->t.R, c->t.S, c->t.sig, c->t.D, c->t.stamp, c->st, c->hr, c->lut,
                  c->arith.nonflush, c->arith.flush, c->arith.offsets, c->dtrace,
                  c->buckets.flop5, c->buckets.turn6};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

int rm_sweep_if_dirty(mccfr_ctx* c) {
  if (!c->sigma_dirty) return MCCFR_OK;
  launch_rm_sweep(c->stream, c->row, c->t, c->slots);
  HIPX(c, hipGetLastError());
  c->sigma_dirty = false;
  return MCCFR_OK;
}

// Slot-major rows <-> action-major planes.
template <class T>
void to_soa(const T* aos, T* soa, uint64_t slots, int row) {
  for (uint64_t s = 0; s < slots; ++s)
    for (int a = 0; a < row; ++a) soa[(uint64_t)a * slots + s] = aos[s * row + a];
}
template <class T>
void to_aos(const T* soa, T* aos, uint64_t slots, int row) {
  for (uint64_t s = 0; s < slots; ++s)
    for (int a = 0; a < row; ++a) aos[s * row + a] = soa[(uint64_t)a * slots + s];
}

}  // namespace

namespace mccfr {

bool make_dev_game(const mccfr_game_cfg& cfg, DevGame& g, const char** why) {
  auto bad = [&](const char* m) {
    if (why) *why = m;
    return false;
  };
  if (cfg.num_players < 2 || cfg.num_players > kMaxPlayers) return bad("num_players must be 2..6");
  if (cfg.num_actions < 1 || cfg.num_actions > MCCFR_MAX_ACTIONS) return bad("num_actions must be 1..16");
  if (cfg.table_log2 < 4 || cfg.table_log2 > 32) return bad("table_log2 must be 4..32");
  if (cfg.history_bits < 1 || cfg.history_bits > 8) return bad("history_bits must be 1..8");
  if (cfg.starting_stack < 0 || cfg.small_blind < 0 || cfg.big_blind < 0) return bad("negative chips");
  if (cfg.starting_stack > 32767) return bad("starting_stack must be <= 32767 chips");
  std::memset(&g, 0, sizeof g);
  g.N = cfg.num_players;
  g.sb = cfg.small_blind;
  g.bb = cfg.big_blind;
  g.num_actions = cfg.num_actions;
  g.row = cfg.num_actions <= 8 ? 8 : 16;
  g.max_raises = cfg.max_raises;
  g.hist_bits = cfg.history_bits;
  g.table_log2 = cfg.table_log2;
  g.use_cfr_plus = cfg.use_cfr_plus;
  g.use_linear_cfr = cfg.use_linear_cfr;
  g.seed_lo = (uint32_t)cfg.seed;
  g.seed_hi = (uint32_t)(cfg.seed >> 32);
  g.tag = 0;
  for (int p = 0; p < kMaxPlayers; ++p) g.stack[p] = p < cfg.num_players ? cfg.starting_stack : 0;
  int have_passive = 0;
  for (int a = 0; a < cfg.num_actions; ++a) {
    const mccfr_action& m = cfg.menu[a];
    if (m.kind > MCCFR_ACT_ALL_IN) return bad("unknown action kind");
    if (m.kind == MCCFR_ACT_POT_FRAC && (m.den == 0 || m.num == 0)) return bad("pot fraction needs num, den > 0");
    g.kind[a] = m.kind;
    g.num[a] = m.num;
    g.den[a] = m.den ? m.den : 1;
    const uint32_t bit = 1u << a;
    switch (m.kind) {
      case MCCFR_ACT_FOLD: g.fold_mask |= bit; break;
      case MCCFR_ACT_CHECK: g.check_mask |= bit; ++have_passive; break;
      case MCCFR_ACT_CALL: g.call_mask |= bit; break;
      case MCCFR_ACT_CHECK_CALL: g.check_mask |= bit; g.call_mask |= bit; ++have_passive; break;
      case MCCFR_ACT_ALL_IN: g.allin_mask |= bit; break;
      default: g.raise_mask |= bit; break;
    }
  }
  if (!have_passive) return bad("menu needs a check or check/call action");
  if (!g.fold_mask) return bad("menu needs a fold action");
  if (!g.allin_mask) return bad("menu needs an all-in action");
  if ((1u << cfg.history_bits) < (uint32_t)cfg.num_actions) return bad("history_bits too small for the menu");
  // passive fallback preference (cfr_gpu.cu:498-508): check, call, check/call, fold,
  // raises in menu order, all-in
  const int order[4] = {MCCFR_ACT_CHECK, MCCFR_ACT_CALL, MCCFR_ACT_CHECK_CALL, MCCFR_ACT_FOLD};
  for (int k : order)
    for (int a = 0; a < cfg.num_actions; ++a)
      if (cfg.menu[a].kind == k) g.passive_order[g.n_passive++] = (uint32_t)a;
  for (int a = 0; a < cfg.num_actions; ++a)
    if (g.raise_mask >> a & 1u) g.passive_order[g.n_passive++] = (uint32_t)a;
  for (int a = 0; a < cfg.num_actions; ++a)
    if (g.allin_mask >> a & 1u) g.passive_order[g.n_passive++] = (uint32_t)a;
  return true;
}

}  // namespace mccfr

extern "C" {

const char* mccfr_last_error(const mccfr_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  std::lock_guard<std::mutex> lk(g_err_mu);
  return g_err.c_str();
}

void mccfr_game_cfg_reference(mccfr_game_cfg* cfg, int num_players, int starting_stack, int small_blind,
                              int big_blind, int use_cfr_plus, int use_linear_cfr) {
  std::memset(cfg, 0, sizeof *cfg);
  cfg->num_players = num_players;
  cfg->starting_stack = starting_stack;
  cfg->small_blind = small_blind;
  cfg->big_blind = big_blind;
  cfg->use_cfr_plus = use_cfr_plus;
  cfg->use_linear_cfr = use_linear_cfr;
  cfg->num_actions = 8;
  cfg->max_raises = 0;
  cfg->history_bits = 3;
  cfg->table_log2 = 26;
  cfg->seed = 42;  // the reference's fixed seed (cfr_gpu.cuh:124)
  const mccfr_action menu[8] = {{MCCFR_ACT_FOLD, 0, 1, 0},     {MCCFR_ACT_CHECK, 0, 1, 0},
                                {MCCFR_ACT_CALL, 0, 1, 0},     {MCCFR_ACT_RAISE_MIN, 0, 1, 0},
                                {MCCFR_ACT_POT_FRAC, 1, 2, 0}, {MCCFR_ACT_POT_FRAC, 2, 3, 0},
                                {MCCFR_ACT_POT_FRAC, 1, 1, 0}, {MCCFR_ACT_ALL_IN, 0, 1, 0}};
  std::memcpy(cfg->menu, menu, sizeof menu);
}

int mccfr_create(const mccfr_game_cfg* cfg, uint32_t max_batch, mccfr_ctx** out) {
  if (!cfg || !out) return fail(nullptr, MCCFR_E_ARG, "null argument");
  *out = nullptr;
  if (max_batch == 0 || max_batch > (1u << 26)) return fail(nullptr, MCCFR_E_ARG, "max_batch must be 1..2^26");
  auto* c = new (std::nothrow) mccfr_ctx();
  if (!c) return fail(nullptr, MCCFR_E_NOMEM, "host allocation failed");
  const char* why = nullptr;
  if (!make_dev_game(*cfg, c->hg, &why)) {
    delete c;
    return fail(nullptr, MCCFR_E_ARG, why);
  }
  c->cfg = *cfg;
  c->row = c->hg.row;
  c->slots = 1ull << cfg->table_log2;
  c->max_batch = max_batch;
  auto bail = [&](int code) {
    std::string m = c->err;
    free_all(c);
    delete c;
    return fail(nullptr, code, m);
  };
#define CHK(call)                                                                                \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      c->err = std::string(#call) + ": " + hipGetErrorString(e_);                               \
      return bail(e_ == hipErrorOutOfMemory ? MCCFR_E_NOMEM : MCCFR_E_HIP);                     \
    }                                                                                           \
  } while (0)
  CHK(hipGetDevice(&c->device));
  {
    int cus = 0;
    CHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    (void)cus;
    c->walk_blocks = 0;  // one lane per traversal: measured fastest on MI355X (persistent grids 5-80% slower)
  }
  CHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  const size_t elems = (size_t)c->slots * c->row;
  CHK(hipMalloc(&c->t.R, elems * sizeof(float)));
  CHK(hipMalloc(&c->t.S, elems * sizeof(float)));
  CHK(hipMalloc(&c->t.sig, elems * sizeof(uint16_t)));
  CHK(hipMalloc(&c->t.D, elems * 2 * sizeof(double)));
  CHK(hipMalloc(&c->t.stamp, (size_t)c->slots * sizeof(uint32_t)));
  CHK(alloc_node_lists(c->t, max_batch));
  CHK(hipMalloc(&c->st, sizeof(LaunchState)));
  CHK(hipMalloc(&c->dg, sizeof(DevGame)));
  CHK(hipMalloc(&c->lut, 2704));
  CHK(hipMalloc(&c->buckets.flop5, kFlopSets));
  CHK(hipMalloc(&c->buckets.turn6, kTurnSets));
  CHK(hipMemsetAsync(c->t.R, 0, elems * sizeof(float), c->stream));
  CHK(hipMemsetAsync(c->t.S, 0, elems * sizeof(float), c->stream));
  CHK(hipMemsetAsync(c->t.sig, 0, elems * sizeof(uint16_t), c->stream));
  CHK(hipMemsetAsync(c->t.D, 0, elems * 2 * sizeof(double), c->stream));
  CHK(hipMemsetAsync(c->t.stamp, 0, (size_t)c->slots * sizeof(uint32_t), c->stream));
  CHK(hipMemsetAsync(c->st, 0, sizeof(LaunchState), c->stream));
  CHK(hipMemcpyAsync(c->dg, &c->hg, sizeof(DevGame), hipMemcpyHostToDevice, c->stream));
  uint8_t lut[2704];
  build_preflop_lut(lut);
  CHK(hipMemcpyAsync(c->lut, lut, sizeof lut, hipMemcpyHostToDevice, c->stream));
  std::vector<uint16_t> nf, fl, off;
  build_arith_tables(nf, fl, off);
  CHK(hipMalloc(&c->arith.nonflush, nf.size() * 2));
  CHK(hipMalloc(&c->arith.flush, fl.size() * 2));
  CHK(hipMalloc(&c->arith.offsets, off.size() * 2));
  CHK(hipMemcpyAsync(c->arith.nonflush, nf.data(), nf.size() * 2, hipMemcpyHostToDevice, c->stream));
  CHK(hipMemcpyAsync(c->arith.flush, fl.data(), fl.size() * 2, hipMemcpyHostToDevice, c->stream));
  CHK(hipMemcpyAsync(c->arith.offsets, off.data(), off.size() * 2, hipMemcpyHostToDevice, c->stream));
  CHK(hipStreamSynchronize(c->stream));
#undef CHK
  *out = c;
  return MCCFR_OK;
}

void mccfr_destroy(mccfr_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
}

int mccfr_generate_hand_table(int32_t* hr, size_t words, int threads) {
  if (!hr || words != kHandTableWords) return fail(nullptr, MCCFR_E_ARG, "hand table needs 32,487,834 words");
  const int n = generate_hand_table(hr, threads);
  if (n != kHandStates) return fail(nullptr, MCCFR_E_TABLE, "hand table generator produced a wrong state count");
  return MCCFR_OK;
}

int mccfr_set_hand_table(mccfr_ctx* c, const int32_t* hr, size_t words) {
  if (!c || !hr) return fail(c, MCCFR_E_ARG, "null argument");
  if (!table_ok(hr, words))
    return fail(c, MCCFR_E_TABLE,
                "incompatible hand table: expected 32,487,834 words with As Ks Qs Js Ts 2c 3d = 7462");
  if (!c->hr) HIPX(c, hipMalloc(&c->hr, words * sizeof(int32_t)));
  HIPX(c, hipMemcpyAsync(c->hr, hr, words * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  launch_build_buckets(c->stream, c->hr, c->buckets);  // flop/turn buckets of every 5-/6-card set
  HIPX(c, hipGetLastError());
  HIPX(c, hipStreamSynchronize(c->stream));
  return MCCFR_OK;
}

int mccfr_load_hand_table(mccfr_ctx* c, const char* path) {
  if (!c || !path) return fail(c, MCCFR_E_ARG, "null argument");
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f.is_open()) return fail(c, MCCFR_E_IO, std::string("cannot open ") + path);
  const size_t bytes = (size_t)f.tellg();
  f.seekg(0);
  std::vector<int32_t> v(bytes / sizeof(int32_t));
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(int32_t)));
  if (!f) return fail(c, MCCFR_E_IO, std::string("cannot read ") + path);
  return mccfr_set_hand_table(c, v.data(), v.size());
}

int mccfr_use_generated_hand_table(mccfr_ctx* c) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  const std::vector<int32_t>* t = generated_table();
  if (!t) return fail(c, MCCFR_E_TABLE, "hand table generation failed");
  return mccfr_set_hand_table(c, t->data(), t->size());
}

int mccfr_preflop_lut(uint8_t* out) {
  if (!out) return fail(nullptr, MCCFR_E_ARG, "null argument");
  build_preflop_lut(out);
  return MCCFR_OK;
}

int mccfr_eval7_device(mccfr_ctx* c, const uint8_t* cards, size_t stride, uint16_t* ranks, size_t n,
                       int evaluator) {
  if (!c || (!cards && n) || (!ranks && n) || stride < 7) return fail(c, MCCFR_E_ARG, "bad eval7 arguments");
  if (n == 0) return MCCFR_OK;
  if (evaluator == 0) {
    if (!c->hr) return fail(c, MCCFR_E_STATE, "no hand table loaded");
    launch_eval7_walk(c->stream, c->hr, cards, stride, ranks, n);
  } else if (evaluator == 1) {
    launch_eval7_arith(c->stream, c->arith, cards, stride, ranks, n);
  } else {
    return fail(c, MCCFR_E_ARG, "evaluator must be 0 (table walk) or 1 (arithmetic)");
  }
  HIPX(c, hipGetLastError());
  return MCCFR_OK;
}

int mccfr_eval7_host(mccfr_ctx* c, const uint8_t* cards, size_t stride, uint16_t* ranks, size_t n,
                     int evaluator) {
  if (!c || (!cards && n) || (!ranks && n) || stride < 7) return fail(c, MCCFR_E_ARG, "bad eval7 arguments");
  if (n == 0) return MCCFR_OK;
  uint8_t* dc = nullptr;
  uint16_t* dr = nullptr;
  HIPX(c, hipMalloc(&dc, n * stride));
  hipError_t e = hipMalloc(&dr, n * sizeof(uint16_t));
  if (e != hipSuccess) {
    (void)hipFree(dc);
    return fail(c, MCCFR_E_NOMEM, "device allocation failed");
  }
  int rc = MCCFR_OK;
  e = hipMemcpyAsync(dc, cards, n * stride, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) rc = mccfr_eval7_device(c, dc, stride, dr, n, evaluator);
  if (e == hipSuccess && rc == MCCFR_OK)
    e = hipMemcpyAsync(ranks, dr, n * sizeof(uint16_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dc);
  (void)hipFree(dr);
  if (e != hipSuccess) return fail(c, MCCFR_E_HIP, hipGetErrorString(e));
  return rc;
}

int mccfr_random_hands_device(mccfr_ctx* c, uint64_t seed, uint32_t tag, uint8_t* d_cards8, size_t n) {
  if (!c || (!d_cards8 && n)) return fail(c, MCCFR_E_ARG, "bad arguments");
  if (n == 0) return MCCFR_OK;
  launch_random_hands(c->stream, seed, tag, d_cards8, n);
  HIPX(c, hipGetLastError());
  return MCCFR_OK;
}

static int record_event(mccfr_ctx* c, int k) {
  if (!c->timing) return MCCFR_OK;
  hipEvent_t e = c->ev[(size_t)(c->ev_next % kTimingRing) * 3 + k];
  HIPX(c, hipEventRecord(e, c->stream));
  return MCCFR_OK;
}

int mccfr_run_launch(mccfr_ctx* c, int update_player, int64_t t_weight, uint64_t first_tid, uint32_t n,
                     mccfr_trace* trace) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (!c->hr) return fail(c, MCCFR_E_STATE, "no hand table loaded");
  if (update_player < 0 || update_player >= c->cfg.num_players) return fail(c, MCCFR_E_ARG, "bad update player");
  if (n == 0 || n > c->max_batch) return fail(c, MCCFR_E_ARG, "n must be 1..max_batch (see mccfr_set_max_batch)");
  int rc = rm_sweep_if_dirty(c);
  if (rc) return rc;
  if (trace && !c->dtrace) HIPX(c, hipMalloc(&c->dtrace, (size_t)c->max_batch * sizeof(mccfr_trace)));
  launch_set_state(c->stream, c->st, update_player, c->next_iter, t_weight, first_tid, n);
  if ((rc = record_event(c, 0))) return rc;
  launch_traverse(c->stream, c->row, c->dg, c->st, c->buckets, c->arith, c->lut, c->t, n, c->walk_blocks,
                  c->scatter_wg, trace ? c->dtrace : nullptr);
  if ((rc = record_event(c, 1))) return rc;
  launch_update(c->stream, c->row, c->dg, c->st, c->t, n);
  if ((rc = record_event(c, 2))) return rc;
  HIPX(c, hipGetLastError());
  if (c->timing) {
    ++c->ev_next;
    c->ev_used = std::min(c->ev_used + 1, kTimingRing);
  }
  c->traversals += n;
  ++c->launches;
  if (trace) {
    HIPX(c, hipMemcpyAsync(trace, c->dtrace, (size_t)n * sizeof(mccfr_trace), hipMemcpyDeviceToHost, c->stream));
    HIPX(c, hipStreamSynchronize(c->stream));
  }
  return MCCFR_OK;
}

int mccfr_train(mccfr_ctx* c, int64_t iters, uint32_t batch) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (!c->hr) return fail(c, MCCFR_E_STATE, "no hand table loaded");
  if (iters <= 0) return MCCFR_OK;
  if (batch == 0 || batch > c->max_batch) return fail(c, MCCFR_E_ARG, "batch must be 1..max_batch");
  int rc = rm_sweep_if_dirty(c);
  if (rc) return rc;
  if (c->gexec && c->graph_batch != batch) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  if (!c->gexec) {
    // one player-update: [traverse -> touched-slot update -> advance launch state]
    hipGraph_t g = nullptr;
    HIPX(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    launch_traverse(c->stream, c->row, c->dg, c->st, c->buckets, c->arith, c->lut, c->t, batch, c->walk_blocks,
                    c->scatter_wg, nullptr);
    launch_update(c->stream, c->row, c->dg, c->st, c->t, batch);
    launch_advance_state(c->stream, c->st, c->cfg.num_players, c->cfg.use_linear_cfr, batch);
    HIPX(c, hipStreamEndCapture(c->stream, &g));
    hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(c, MCCFR_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    c->graph_batch = batch;
  }
  const int64_t it0 = c->next_iter;
  launch_set_state(c->stream, c->st, 0, it0, c->cfg.use_linear_cfr ? it0 : 1, c->traversals, batch);
  HIPX(c, hipGetLastError());
  const int N = c->cfg.num_players;
  for (int64_t it = 0; it < iters; ++it)
    for (int p = 0; p < N; ++p) HIPX(c, hipGraphLaunch(c->gexec, c->stream));
  c->traversals += (uint64_t)iters * N * batch;
  c->launches += (uint64_t)iters * N;
  c->next_iter += iters;
  HIPX(c, hipStreamSynchronize(c->stream));
  return MCCFR_OK;
}

int mccfr_set_max_batch(mccfr_ctx* c, uint32_t max_batch) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (max_batch == 0 || max_batch > (1u << 26)) return fail(c, MCCFR_E_ARG, "max_batch must be 1..2^26");
  HIPX(c, hipStreamSynchronize(c->stream));
  if (max_batch != c->t.lanes) {
    free_node_lists(c->t);
    const hipError_t e = alloc_node_lists(c->t, max_batch);
    if (e != hipSuccess) {
      free_node_lists(c->t);
      c->t.lanes = 0;
      c->max_batch = 0;
      return fail(c, MCCFR_E_NOMEM, std::string("node lists: ") + hipGetErrorString(e));
    }
  }
  if (c->dtrace) {
    (void)hipFree(c->dtrace);
    c->dtrace = nullptr;
  }
  if (c->gexec) {  // the graph holds the old scratch pointers
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  c->max_batch = max_batch;
  return MCCFR_OK;
}

static constexpr uint64_t kCheckpointMagic = 0x43465247505500ull;  // cfr_gpu.cu:1480

int mccfr_save_checkpoint(mccfr_ctx* c, const char* path) {
  if (!c || !path) return fail(c, MCCFR_E_ARG, "null argument");
  if (c->slots > 0xFFFFFFFFull) return fail(c, MCCFR_E_ARG, "table too large for the .ckpt header");
  const size_t n = (size_t)c->slots * c->row;
  std::vector<float> R(n), S(n);
  int rc = mccfr_download_tables(c, R.data(), S.data(), nullptr, n, 1);
  if (rc) return rc;
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(c, MCCFR_E_IO, std::string("cannot open ") + path);
  const uint64_t magic = kCheckpointMagic;
  const uint32_t ts = (uint32_t)c->slots, na = (uint32_t)c->row;
  f.write(reinterpret_cast<const char*>(&magic), 8);
  f.write(reinterpret_cast<const char*>(&ts), 4);
  f.write(reinterpret_cast<const char*>(&na), 4);
  f.write(reinterpret_cast<const char*>(R.data()), (std::streamsize)(n * sizeof(float)));
  f.write(reinterpret_cast<const char*>(S.data()), (std::streamsize)(n * sizeof(float)));
  if (!f) return fail(c, MCCFR_E_IO, std::string("write failed: ") + path);
  return MCCFR_OK;
}

int mccfr_load_checkpoint(mccfr_ctx* c, const char* path) {
  if (!c || !path) return fail(c, MCCFR_E_ARG, "null argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(c, MCCFR_E_IO, std::string("cannot open ") + path);
  uint64_t magic = 0;
  uint32_t ts = 0, na = 0;
  f.read(reinterpret_cast<char*>(&magic), 8);
  f.read(reinterpret_cast<char*>(&ts), 4);
  f.read(reinterpret_cast<char*>(&na), 4);
  if (!f) return fail(c, MCCFR_E_IO, std::string("cannot read header of ") + path);
  if (magic != kCheckpointMagic) return fail(c, MCCFR_E_FORMAT, std::string(path) + " is not a checkpoint");
  if ((uint64_t)ts != c->slots || na != (uint32_t)c->row)
    return fail(c, MCCFR_E_FORMAT, std::string(path) + " was saved with table_size=" + std::to_string(ts) +
                                       " num_actions=" + std::to_string(na) + ", this context has " +
                                       std::to_string(c->slots) + " and " + std::to_string(c->row));
  const size_t n = (size_t)ts * na;
  std::vector<float> R(n), S(n);
  f.read(reinterpret_cast<char*>(R.data()), (std::streamsize)(n * sizeof(float)));
  f.read(reinterpret_cast<char*>(S.data()), (std::streamsize)(n * sizeof(float)));
  if (!f) return fail(c, MCCFR_E_IO, std::string("truncated checkpoint ") + path);
  return mccfr_upload_tables(c, R.data(), S.data(), n, 1);
}

int mccfr_set_walk_blocks(mccfr_ctx* c, uint32_t blocks) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  c->walk_blocks = blocks;
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  return MCCFR_OK;
}

int mccfr_set_scatter_wg(mccfr_ctx* c, uint32_t wg) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (wg != 256 && wg != 512 && wg != 1024) return fail(c, MCCFR_E_ARG, "scatter workgroup must be 256, 512 or 1024");
  c->scatter_wg = wg;
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  return MCCFR_OK;
}

int mccfr_get_walk_blocks(const mccfr_ctx* c, uint32_t* blocks) {
  if (!c || !blocks) return fail(nullptr, MCCFR_E_ARG, "null argument");
  *blocks = c->walk_blocks;
  return MCCFR_OK;
}

int mccfr_randomize_tables(mccfr_ctx* c, uint64_t seed) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  launch_randomize(c->stream, c->row, c->cfg.num_actions, c->t, c->slots, seed);
  HIPX(c, hipGetLastError());
  c->sigma_dirty = true;
  return rm_sweep_if_dirty(c);
}

int mccfr_sync(mccfr_ctx* c) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  HIPX(c, hipStreamSynchronize(c->stream));
  return MCCFR_OK;
}

int mccfr_table_shape(const mccfr_ctx* c, uint64_t* slots, uint32_t* row) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (slots) *slots = c->slots;
  if (row) *row = (uint32_t)c->row;
  return MCCFR_OK;
}

int mccfr_download_tables(mccfr_ctx* c, float* R, float* S, uint16_t* sig, size_t elems, int layout) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  const size_t want = (size_t)c->slots * c->row;
  if (elems != want) return fail(c, MCCFR_E_ARG, "elems must equal slots * row_actions");
  if (layout != 0 && layout != 1) return fail(c, MCCFR_E_ARG, "layout must be 0 (rows) or 1 (planes)");
  int rc = rm_sweep_if_dirty(c);
  if (rc) return rc;
  HIPX(c, hipStreamSynchronize(c->stream));
  auto get = [&](auto* host, const auto* dev) -> int {
    using T = std::remove_pointer_t<decltype(host)>;
    if (!host) return MCCFR_OK;
    if (layout == 0) {
      HIPX(c, hipMemcpy(host, dev, want * sizeof(T), hipMemcpyDeviceToHost));
    } else {
      std::vector<T> tmp(want);
      HIPX(c, hipMemcpy(tmp.data(), dev, want * sizeof(T), hipMemcpyDeviceToHost));
      to_soa(tmp.data(), host, c->slots, c->row);
    }
    return MCCFR_OK;
  };
  if ((rc = get(R, c->t.R))) return rc;
  if ((rc = get(S, c->t.S))) return rc;
  if ((rc = get(sig, c->t.sig))) return rc;
  return MCCFR_OK;
}

int mccfr_upload_tables(mccfr_ctx* c, const float* R, const float* S, size_t elems, int layout) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  const size_t want = (size_t)c->slots * c->row;
  if (elems != want) return fail(c, MCCFR_E_ARG, "elems must equal slots * row_actions");
  if (layout != 0 && layout != 1) return fail(c, MCCFR_E_ARG, "layout must be 0 (rows) or 1 (planes)");
  HIPX(c, hipStreamSynchronize(c->stream));
  auto put = [&](const float* host, float* dev) -> int {
    if (!host) return MCCFR_OK;
    if (layout == 0) {
      HIPX(c, hipMemcpy(dev, host, want * sizeof(float), hipMemcpyHostToDevice));
    } else {
      std::vector<float> tmp(want);
      to_aos(host, tmp.data(), c->slots, c->row);
      HIPX(c, hipMemcpy(dev, tmp.data(), want * sizeof(float), hipMemcpyHostToDevice));
    }
    return MCCFR_OK;
  };
  int rc;
  if ((rc = put(R, c->t.R))) return rc;
  if ((rc = put(S, c->t.S))) return rc;
  c->sigma_dirty = true;  // the next launch regret-matches the whole table first (cfr_gpu.cu:1290)
  return MCCFR_OK;
}

int mccfr_normalized_strategy(mccfr_ctx* c, float* out, size_t elems, int layout) {
  if (!c || !out) return fail(c, MCCFR_E_ARG, "null argument");
  const size_t want = (size_t)c->slots * c->row;
  if (elems != want) return fail(c, MCCFR_E_ARG, "elems must equal slots * row_actions");
  if (layout != 0 && layout != 1) return fail(c, MCCFR_E_ARG, "layout must be 0 (rows) or 1 (planes)");
  float* d = nullptr;
  HIPX(c, hipMalloc(&d, want * sizeof(float)));
  launch_normalize(c->stream, c->row, c->t.S, d, c->slots);
  std::vector<float> tmp(layout ? want : 0);
  hipError_t e = hipMemcpyAsync(layout ? tmp.data() : out, d, want * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, MCCFR_E_HIP, hipGetErrorString(e));
  if (layout) to_soa(tmp.data(), out, c->slots, c->row);
  return MCCFR_OK;
}

int mccfr_counters(const mccfr_ctx* c, uint64_t* traversals, uint64_t* launches, int64_t* next_iteration) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (traversals) *traversals = c->traversals;
  if (launches) *launches = c->launches;
  if (next_iteration) *next_iteration = c->next_iter;
  return MCCFR_OK;
}

int mccfr_set_timing(mccfr_ctx* c, int enabled) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (enabled && c->ev.empty()) {
    c->ev.resize((size_t)kTimingRing * 3, nullptr);
    for (auto& e : c->ev) HIPX(c, hipEventCreate(&e));
  }
  c->timing = enabled != 0;
  c->ev_next = 0;
  c->ev_used = 0;
  return MCCFR_OK;
}

int mccfr_kernel_times(mccfr_ctx* c, float* trav_ms, float* upd_ms, uint32_t n) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (n > c->ev_used) return fail(c, MCCFR_E_ARG, "fewer timed launches recorded than requested");
  HIPX(c, hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t k = (c->ev_next - n + i) % kTimingRing;
    float a = 0.f, b = 0.f;
    HIPX(c, hipEventElapsedTime(&a, c->ev[k * 3 + 0], c->ev[k * 3 + 1]));
    HIPX(c, hipEventElapsedTime(&b, c->ev[k * 3 + 1], c->ev[k * 3 + 2]));
    if (trav_ms) trav_ms[i] = a;
    if (upd_ms) upd_ms[i] = b;
  }
  return MCCFR_OK;
}

int mccfr_copy_probe(mccfr_ctx* c, size_t bytes, int reps, float* ms_per_copy) {
  if (!c || bytes < 16 || reps < 1) return fail(c, MCCFR_E_ARG, "bad copy probe arguments");
  bytes &= ~(size_t)15;
  void *a = nullptr, *b = nullptr;
  HIPX(c, hipMalloc(&a, bytes));
  hipError_t e = hipMalloc(&b, bytes);
  if (e != hipSuccess) {
    (void)hipFree(a);
    return fail(c, MCCFR_E_NOMEM, "device allocation failed");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipMemsetAsync(a, 1, bytes, c->stream);
  launch_copy(c->stream, (const uint4*)a, (uint4*)b, bytes / 16);  // warm-up
  (void)hipEventRecord(e0, c->stream);
  for (int r = 0; r < reps; ++r) launch_copy(c->stream, (const uint4*)a, (uint4*)b, bytes / 16);
  (void)hipEventRecord(e1, c->stream);
  e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(c, MCCFR_E_HIP, hipGetErrorString(e));
  if (ms_per_copy) *ms_per_copy = ms / reps;
  return MCCFR_OK;
}

// ---- device-memory plumbing for hosts without their own HIP bindings (tests, bench)
int mccfr_device_alloc(mccfr_ctx* c, size_t bytes, void** out) {
  if (!c || !out) return fail(c, MCCFR_E_ARG, "null argument");
  HIPX(c, hipMalloc(out, bytes));
  return MCCFR_OK;
}
int mccfr_device_free(mccfr_ctx* c, void* p) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  HIPX(c, hipStreamSynchronize(c->stream));
  HIPX(c, hipFree(p));
  return MCCFR_OK;
}
int mccfr_memcpy(mccfr_ctx* c, void* dst, const void* src, size_t bytes, int kind) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  HIPX(c, hipMemcpyAsync(dst, src, bytes, k, c->stream));
  HIPX(c, hipStreamSynchronize(c->stream));
  return MCCFR_OK;
}

}  // extern "C"
