// Host runtime behind the C ABI (include/mccfr/mccfr.h): contexts, tenants,
// device buffers, table upload/download, batched launch sequencing, the
// hipGraph training loop and the launch-shape autotuner.  Replaces
// GPUCFRTrainer's host side (cfr_gpu.cu:1064-1538).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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
#include "runtime.hpp"

using namespace mccfr;

namespace {

std::string g_err;
std::mutex g_err_mu;

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

// Hand-table check (hand_eval.cpp:59-86: size, then As Ks Qs Js Ts 2c 3d must
// rank 7462), plus a range check the reference does not make: every word must
// lie in [0, words - 53].  A walk reads hr[p + c] for card codes c in 1..52
// with p the previous word, so with every word in range no walk over this table
// -- the rank walks, the bucket builder's 5- and 6-card walks that decode a
// state word, the harness's walks -- can read outside it, whatever its
// contents.  Valid tables pass: state words are state bases (<= words - 53) and
// rank words are below 0xA000.  On failure *why names the check.
int table_check(const int32_t* hr, size_t words, const char** why) {
  if (words != kHandTableWords) {
    *why = "expected 32,487,834 words";
    return MCCFR_E_TABLE;
  }
  const uint32_t hi = (uint32_t)(words - 53);
  // one linear pass; an out-of-range word sets a bit of `bad`
  uint32_t bad = 0;
  for (size_t i = 0; i < words; ++i) bad |= (uint32_t)hr[i] > hi;
  if (bad) {
    *why = "a word lies outside [0, 32,487,781]";
    return MCCFR_E_TABLE;
  }
  const uint8_t royal[7] = {51, 50, 49, 48, 47, 0, 14};  // As Ks Qs Js Ts 2c 3d
  int p = hr[53 + card_rm(royal[0])];
  for (int i = 1; i < 6; ++i) p = hr[p + card_rm(royal[i])];
  if (decode_packed((uint32_t)hr[p + card_rm(royal[6])] & 0xFFFFu) != 7462u) {
    *why = "As Ks Qs Js Ts 2c 3d does not rank 7462";
    return MCCFR_E_TABLE;
  }
  return MCCFR_OK;
}

// Reads a hand-table file into v after checking its size (no allocation for a
// file of the wrong size), then checks its contents.
int read_table_file(const char* path, std::vector<int32_t>& v, std::string& msg) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f.is_open()) {
    msg = std::string("cannot open ") + path;
    return MCCFR_E_IO;
  }
  const std::streamoff bytes = f.tellg();
  if (bytes != (std::streamoff)(kHandTableWords * sizeof(int32_t))) {
    msg = std::string("incompatible hand table ") + path + ": expected 129,951,336 bytes";
    return MCCFR_E_TABLE;
  }
  f.seekg(0);
  try {
    v.resize(kHandTableWords);
  } catch (...) {
    msg = "no host memory for the hand table";
    return MCCFR_E_NOMEM;
  }
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(int32_t)));
  if (!f) {
    msg = std::string("cannot read ") + path;
    return MCCFR_E_IO;
  }
  const char* why = "";
  const int rc = table_check(v.data(), v.size(), &why);
  if (rc != MCCFR_OK) msg = std::string("incompatible hand table ") + path + ": " + why;
  return rc;
}

uint32_t round_up(uint32_t x, uint32_t a) { return (x + a - 1) / a * a; }

// Launch scratch for `cap` lanes (a multiple of kSegAlign): one arena holding
// every buffer of DevScratch at 256-byte aligned offsets.  The byte counts of
// the parts, in arena order (scratch_parts), size both the context's arena
// (one hipMalloc) and the allocation-free entry point's static arena.
constexpr int kScratchParts = 28;
struct ScratchParts {
  size_t b[kScratchParts];
};
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr ScratchParts scratch_parts(uint32_t cap) {
  const size_t L = cap;
  const size_t recs = L * kMaxDepth;
  const size_t max_tiles = (recs + kSortTile - 1) / kSortTile;
  const size_t max_long = recs / (kShortRun + 1) + 1;
  const size_t max_pieces = recs / kPiece + max_long;
  const size_t max_waves = recs / 64 + 1;
  return ScratchParts{{recs * 4, recs * 4, L, L * 4,                              // node_slot node_av node_count node_u
                       (size_t)kMaxPlayers * L * 4, (size_t)kMaxPlayers * L * 2,  // root_bkt root_rank
                       (size_t)kMaxPlayers * L * 4, (size_t)kMaxPlayers * L * 2,  // root_tidx sd_inv
                       L, L, ((L + kEmitLanes - 1) / kEmitLanes) * 4,            // sd_state root_dealer blk_cnt
                       recs * 4, recs * 8, recs * 4, recs * 8,                    // key[0] val[0] key[1] val[1]
                       (size_t)kSortBins * max_tiles * 4, (size_t)kSortBins * 4,  // hist bin_total
                       max_waves * 64 * 16, max_waves * 4, max_waves * 4 * 8,     // runs n_runs cand
                       max_waves * 4, max_waves + kHeadPad, head1_bytes(cap),     // n_cand head0 head1
                       (size_t)kMaxSegs * sizeof(SegCtx), max_long * 16,          // seg_ctx long_runs
                       max_pieces * 16, max_pieces * 32 * 8, sizeof(LaunchCounters)}};  // piece partial cnt
}
constexpr size_t scratch_bytes(uint32_t cap) {
  const ScratchParts z = scratch_parts(cap);
  size_t t = 0;
  for (int i = 0; i < kScratchParts; ++i) t += align256(z.b[i]);
  return t;
}

// Points the buffers of `s` into the arena at `base` (scratch_bytes(cap) bytes).
void carve_scratch(DevScratch& s, char* base, uint32_t cap) {
  std::memset(&s, 0, sizeof s);
  const ScratchParts z = scratch_parts(cap);
  char* p = base;
  int i = 0;
  auto take = [&](auto*& ptr) {
    ptr = reinterpret_cast<std::remove_reference_t<decltype(ptr)>>(p);
    p += align256(z.b[i++]);
  };
  take(s.node_slot), take(s.node_av), take(s.node_count), take(s.node_u), take(s.root_bkt), take(s.root_rank);
  take(s.root_tidx), take(s.sd_inv), take(s.sd_state), take(s.root_dealer), take(s.blk_cnt);
  take(s.key[0]), take(s.val[0]), take(s.key[1]), take(s.val[1]), take(s.hist), take(s.bin_total);
  take(s.runs), take(s.n_runs), take(s.cand), take(s.n_cand), take(s.head0), take(s.head1);
  take(s.seg_ctx), take(s.long_runs), take(s.piece), take(s.partial), take(s.cnt);
  s.cap_lanes = cap;
  s.max_tiles = (uint32_t)(((size_t)cap * kMaxDepth + kSortTile - 1) / kSortTile);
}

hipError_t alloc_scratch(DevScratch& s, uint32_t cap) {
  std::memset(&s, 0, sizeof s);
  char* base = nullptr;
  hipError_t e = hipMalloc(&base, scratch_bytes(cap));
  if (e != hipSuccess) return e;
  carve_scratch(s, base, cap);
  // the walk adds each emission block's record count to blk_cnt, and the
  // emission clears what it read: zero at the start
  e = hipMemset(s.blk_cnt, 0, ((size_t)(cap + kEmitLanes - 1) / kEmitLanes) * 4);
  if (e != hipSuccess) return e;
  return hipMemset(s.cnt, 0, sizeof(LaunchCounters));
}

// The arena starts at node_slot.
void free_scratch(DevScratch& s) {
  if (s.node_slot) (void)hipFree(s.node_slot);
  s.node_slot = nullptr;
}

// Launch scratch of mccfr_run_launch_noalloc: a static arena for launches of
// up to kNoAllocLanes traversals (scratch_bytes(16384) = about 24 MB), held by
// every device the library runs on until the process ends.  In host code the
// name `scratch` is a host-side stand-in, not a device address: the device
// address is taken once per device (noalloc_scratch) and only that is used.
constexpr uint32_t kNoAllocLanes = 16384;
constexpr int kMaxNoAllocDevices = 64;
std::mutex g_noalloc_mu;
char* g_noalloc_ptr[kMaxNoAllocDevices] = {};

void drop_graph(mccfr_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
  c->graph_key.clear();
}

void free_all(mccfr_ctx* c) {
  drop_graph(c);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->stage_ev)
    if (e) (void)hipEventDestroy(e);
  free_scratch(c->sc);
  for (auto& t : c->tenants)
    if (t.used && t.owned) (void)hipFree(t.owned);
  void* ptrs[] = {c->d_tenants, c->d_segs, c->hr, c->lut, c->arith.nonflush, c->arith.flush, c->arith.offsets,
                  c->dtrace, c->buckets.flop5, c->buckets.turn6, c->eval_bk.flop5, c->eval_bk.turn6,
                  c->es_stack, c->es_rd, c->es_gen};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

Tenant* get_tenant(mccfr_ctx* c, uint32_t id) {
  if (!c || id >= c->tenants.size() || !c->tenants[id].used) return nullptr;
  return &c->tenants[id];
}

int rm_sweep_if_dirty(mccfr_ctx* c, Tenant& t) {
  if (!t.sigma_dirty) return MCCFR_OK;
  launch_rm_sweep(c->stream, t.row, t.hg.R, t.hg.sig, t.slots);
  HIPX(c, hipGetLastError());
  t.sigma_dirty = false;
  return MCCFR_OK;
}

// Frames of the external-sampling walk for a launch of `lanes` lanes, a
// regret-delta row for every record the sort input holds, and the generation
// walk's buffers (all grown, never shrunk; a captured graph holding old
// buffers is dropped).
int ensure_es_stack(mccfr_ctx* c, int row, uint32_t lanes) {
  const auto grow = [&](auto*& buf, size_t& have, size_t want, const char* what) -> int {
    if (want <= have) return MCCFR_OK;
    HIPX(c, hipStreamSynchronize(c->stream));
    drop_graph(c);
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    have = 0;
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) return fail(c, MCCFR_E_NOMEM, std::string(what) + hipGetErrorString(e));
    buf = static_cast<std::remove_reference_t<decltype(buf)>>(p);
    have = want;
    return MCCFR_OK;
  };
  int rc = grow(c->es_stack, c->es_stack_bytes, es_stack_bytes(row, lanes), "external-sampling frames: ");
  if (!rc) {
    const size_t rows = (size_t)c->sc.cap_lanes * kMaxDepth;  // the sort input's records
    rc = grow(c->es_rd, c->es_rd_bytes, rows * (size_t)row * sizeof(float), "external-sampling delta rows: ");
  }
  if (!rc && c->es_gen_cap != c->sc.cap_lanes) {
    // nodes and tasks per generation for the sort input's records
    rc = grow(c->es_gen, c->es_gen_size, es_gen_bytes(c->sc.cap_lanes), "external-sampling generations: ");
    if (rc) {  // the old carving may point into freed memory: forget it
      c->es_gen_bufs = EsGenBufs{};
      c->sc.gen = EsGenBufs{};
      c->es_gen_cap = 0;
    } else {
      HIPX(c, hipStreamSynchronize(c->stream));
      drop_graph(c);  // a graph holds the old carving
      carve_es_gen(c->es_gen_bufs, c->es_gen, c->sc.cap_lanes);
      c->es_gen_cap = c->sc.cap_lanes;
    }
  }
  c->sc.es_stack = c->es_stack;
  c->sc.es_rd = c->es_rd;
  if (c->es_gen_cap == c->sc.cap_lanes) c->sc.gen = c->es_gen_bufs;
  return rc;
}

// Waits for the context's work and reports (and clears) an external-sampling
// launch whose records overflowed the sort input (its tables were left as
// they were).
int check_es(mccfr_ctx* c) {
  uint32_t flag = 0;
  HIPX(c, hipMemcpyAsync(&flag, &c->sc.cnt->es_overflow, sizeof flag, hipMemcpyDeviceToHost, c->stream));
  HIPX(c, hipStreamSynchronize(c->stream));
  if (!flag) return MCCFR_OK;
  const uint32_t zero = 0;
  HIPX(c, hipMemcpy(&c->sc.cnt->es_overflow, &zero, sizeof zero, hipMemcpyHostToDevice));
  return fail(c, MCCFR_E_STATE,
              "an external-sampling launch produced more records than its scratch holds (20 per lane of "
              "max_batch): it updated no table; use fewer traversals per launch or a larger max_batch");
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

// Checked, laid-out segments of one launch.
struct Prepared {
  std::vector<DevSeg> segs;
  LaunchCfg lc{};
  uint32_t key_bits = 0;
  uint32_t sort_bits = 0;  // the low key bits the radix passes order by
  int row = 8;
};

int prepare(mccfr_ctx* c, const HostSeg* hs, uint32_t n, const DevScratch& sc, Prepared& p) {
  if (!c->hr) return fail(c, MCCFR_E_STATE, "no hand table loaded");
  if (n == 0 || n > kMaxSegs) return fail(c, MCCFR_E_ARG, "a launch needs 1..4096 segments");
  p.segs.resize(n);
  std::vector<uint8_t> seen(c->tenants.size(), 0);
  uint32_t lanes = 0, slot_bits = 0;
  for (uint32_t i = 0; i < n; ++i) {
    Tenant* t = get_tenant(c, hs[i].tenant);
    if (!t) return fail(c, MCCFR_E_ARG, "segment " + std::to_string(i) + " names no tenant");
    if (seen[hs[i].tenant]++) return fail(c, MCCFR_E_ARG, "a tenant appears in two segments of one launch");
    if (i == 0) p.row = t->row;
    if (t->row != p.row) return fail(c, MCCFR_E_ARG, "segments of one launch need the same row width");
    if (i == 0) p.lc.es = t->cfg.sampling == MCCFR_SAMPLE_EXTERNAL;
    if ((t->cfg.sampling == MCCFR_SAMPLE_EXTERNAL) != (p.lc.es != 0))
      return fail(c, MCCFR_E_ARG, "segments of one launch need the same sampling");
    if (hs[i].n == 0) return fail(c, MCCFR_E_ARG, "segment with no traversals");
    if (hs[i].update_player < 0 || hs[i].update_player >= t->cfg.num_players)
      return fail(c, MCCFR_E_ARG, "bad update player");
    const uint32_t padded = round_up(hs[i].n, kSegAlign);
    if ((uint64_t)lanes + padded > sc.cap_lanes)
      return fail(c, MCCFR_E_ARG,
                  "launch needs more lanes than its scratch holds (segments are padded to 256; see mccfr_set_max_batch)");
    DevSeg& d = p.segs[i];
    std::memset(&d, 0, sizeof d);
    d.tenant = hs[i].tenant;
    d.lane0 = lanes;
    d.n = hs[i].n;
    d.update_player = hs[i].update_player;
    d.iteration = hs[i].iteration;
    d.t_weight = hs[i].t_weight;
    d.first_tid = hs[i].first_tid;
    d.upd_idx = hs[i].upd_idx;
    lanes += padded;
    slot_bits = std::max(slot_bits, (uint32_t)t->cfg.table_log2);
  }
  const uint32_t seg_bits = n > 1 ? 32u - (uint32_t)__builtin_clz(n - 1) : 0u;
  // external sampling sorts one bit more: its null records take the top key
  const uint32_t es_bit = p.lc.es ? 1u : 0u;
  if (slot_bits + seg_bits + es_bit > 32)
    return fail(c, MCCFR_E_ARG, "too many segments for the tables' slot bits (slot bits + segment bits > 32)");
  p.lc.n_segs = n;
  p.lc.lanes = lanes;
  p.lc.slot_bits = slot_bits;
  p.lc.cap_lanes = sc.cap_lanes;
  p.key_bits = slot_bits + seg_bits + es_bit;
  p.lc.null_key = p.key_bits == 32 ? 0xFFFFFFFFu : (1u << p.key_bits) - 1u;
  // An outcome-sampling launch emits its records in lane order, so in segment
  // order, and has no null records: stable passes over the slot bits alone
  // leave equal slots of different segments adjacent but apart, in segment
  // order, so each run of equal keys is still one slot of one tenant with its
  // records in emission order.  External-sampling launches of several
  // segments fill reservation tails with null keys, which must sort last.
  p.sort_bits = p.lc.es ? p.key_bits : slot_bits;
  return MCCFR_OK;
}

// Copies the segment table to the device through a pinned staging slot.
int stage_segments(mccfr_ctx* c, const std::vector<DevSeg>& segs) {
  const uint32_t k = c->stage_next++ % kStageRing;
  HIPX(c, hipEventSynchronize(c->stage_ev[k]));  // the slot's previous copy has left it
  DevSeg* h = c->h_stage + (size_t)k * kMaxSegs;
  std::memcpy(h, segs.data(), segs.size() * sizeof(DevSeg));
  HIPX(c, hipMemcpyAsync(c->d_segs, h, segs.size() * sizeof(DevSeg), hipMemcpyHostToDevice, c->stream));
  HIPX(c, hipEventRecord(c->stage_ev[k], c->stream));
  return MCCFR_OK;
}

hipEvent_t timing_event(mccfr_ctx* c, int k) {
  return c->timing ? c->ev[(size_t)(c->ev_next % kTimingRing) * 4 + k] : nullptr;
}

int record_event(mccfr_ctx* c, int k) {
  if (!c->timing) return MCCFR_OK;
  HIPX(c, hipEventRecord(timing_event(c, k), c->stream));
  return MCCFR_OK;
}

DealInputs deal_inputs(const mccfr_ctx* c) {
  DealInputs in;
  in.bt = c->buckets;
  in.at = c->arith;
  in.lut = c->lut;
  return in;
}

// The host mirror of advance_segment (device_common.hpp).
void advance(const Tenant& t, HostSeg& s) {
  s.first_tid += s.n;
  if (++s.upd_idx >= (uint32_t)t.hg.n_upd) {
    s.upd_idx = 0;
    s.iteration += 1;
  }
  s.update_player = t.hg.upd_seats[s.upd_idx];
  s.t_weight = t.cfg.use_linear_cfr ? s.iteration : 1;
}

HostSeg next_of(uint32_t id, const Tenant& t, uint32_t n) {
  HostSeg s;
  s.tenant = id;
  s.n = n;
  s.upd_idx = t.upd_idx;
  s.update_player = t.hg.upd_seats[t.upd_idx];
  s.iteration = t.next_iter;
  s.t_weight = t.cfg.use_linear_cfr ? t.next_iter : 1;
  s.first_tid = t.traversals;
  return s;
}

}  // namespace

// The allocation-free launch's arena (see kNoAllocLanes), at file scope.
static __device__ __align__(256) unsigned char scratch[scratch_bytes(kNoAllocLanes)];

namespace mccfr {

int fail(mccfr_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = msg;
  return code;
}

bool make_dev_game(const mccfr_game_cfg& cfg, const mccfr_root_state* root, DevGame& g, const char** why) {
  auto bad = [&](const char* m) {
    if (why) *why = m;
    return false;
  };
  if (cfg.num_players < 2 || cfg.num_players > kMaxPlayers) return bad("num_players must be 2..6");
  if (cfg.num_actions < 1 || cfg.num_actions > MCCFR_MAX_ACTIONS) return bad("num_actions must be 1..16");
  if (cfg.table_log2 < 4 || cfg.table_log2 > 31) return bad("table_log2 must be 4..31");
  if (cfg.history_bits < 1 || cfg.history_bits > 8) return bad("history_bits must be 1..8");
  if (cfg.starting_stack < 0 || cfg.small_blind < 0 || cfg.big_blind < 0) return bad("negative chips");
  if (cfg.starting_stack > 32767) return bad("starting_stack must be <= 32767 chips");
  std::memset(&g, 0, sizeof g);
  const int N = cfg.num_players;
  g.N = N;
  g.sb = cfg.small_blind;
  g.bb = cfg.big_blind;
  g.num_actions = cfg.num_actions;
  g.row = cfg.num_actions <= 8 ? 8 : 16;
  g.max_raises = cfg.max_raises;
  g.hist_bits = cfg.history_bits;
  g.table_log2 = cfg.table_log2;
  g.use_cfr_plus = cfg.use_cfr_plus;
  g.use_linear_cfr = cfg.use_linear_cfr;
  if (cfg.sampling != MCCFR_SAMPLE_OUTCOME && cfg.sampling != MCCFR_SAMPLE_EXTERNAL)
    return bad("sampling must be MCCFR_SAMPLE_OUTCOME or MCCFR_SAMPLE_EXTERNAL");
  g.sampling = cfg.sampling;
  g.seed_lo = (uint32_t)cfg.seed;
  g.seed_hi = (uint32_t)(cfg.seed >> 32);
  g.tag = 0;
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
  for (int a = cfg.num_actions; a < MCCFR_MAX_ACTIONS; ++a) g.den[a] = 1;
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

  if (!root || root->street == 0) {  // the reference's preflop root
    for (int p = 0; p < N; ++p) g.stack[p] = g.root_stack[p] = cfg.starting_stack;
    g.n_board = 0;
    g.deck_size = 52;
    g.n_deal = (uint8_t)(2 * N + 5);
    g.walk_k0 = (uint32_t)(2 * N + 6);
    for (int p = 0; p < N; ++p) g.upd_seats[p] = (uint8_t)p;
    g.n_upd = N;
    return true;
  }
  if (root->street < 1 || root->street > 3) return bad("root street must be 0 (preflop), 1, 2 or 3");
  if (root->dealer < 0 || root->dealer >= N) return bad("root dealer must be a seat");
  const uint32_t all = (1u << N) - 1u;
  if (root->active_mask & ~all) return bad("root active_mask names seats beyond num_players");
  if (__builtin_popcount(root->active_mask) < 2) return bad("a root needs at least two active seats");
  const int nb = root->street + 2;
  uint64_t bm = 0;
  for (int k = 0; k < nb; ++k) {
    if (root->board[k] >= 52) return bad("board card codes must be 0..51");
    if (bm >> root->board[k] & 1ull) return bad("board cards must be distinct");
    bm |= 1ull << root->board[k];
    g.board[k] = root->board[k];
  }
  int pot = 0;
  for (int p = 0; p < N; ++p) {
    const int st = root->stack[p], inv = root->invested[p];
    if (st < 0 || inv < 0) return bad("negative root stack or investment");
    if (st + inv > 32767) return bad("root stack + investment must be <= 32767 chips");
    g.root_stack[p] = st;
    g.stack[p] = st + inv;
    pot += inv;
    const bool act = root->active_mask >> p & 1u;
    if (!act) g.root_folded |= 1u << p;
    if (act && st == 0) g.root_allin |= 1u << p;
  }
  g.root_street = root->street;
  g.root_dealer = root->dealer;
  g.root_pot = pot;
  g.root_ab = root->history & kHistoryMask;
  g.board_lo = (uint32_t)bm;
  g.board_hi = (uint32_t)(bm >> 32);
  g.n_board = (uint8_t)nb;
  g.deck_size = (uint8_t)(52 - nb);
  g.n_deal = (uint8_t)(2 * N + 5 - nb);
  g.walk_k0 = (uint32_t)(2 * N + 5 - nb);
  g.n_upd = 0;
  for (int p = 0; p < N; ++p)
    if (root->active_mask >> p & 1u) g.upd_seats[g.n_upd++] = (uint8_t)p;
  return true;
}

uint64_t tenant_bytes(const mccfr_game_cfg& cfg) {
  const uint64_t row = cfg.num_actions <= 8 ? 8 : 16;
  return (1ull << cfg.table_log2) * row * (4 + 4 + 2);
}

int tenant_create(mccfr_ctx* c, const mccfr_game_cfg* cfg, const mccfr_root_state* root, void* mem, uint32_t* id) {
  if (!cfg || !id) return fail(c, MCCFR_E_ARG, "null argument");
  Tenant t;
  const char* why = nullptr;
  if (!make_dev_game(*cfg, root, t.hg, &why)) return fail(c, MCCFR_E_ARG, why);
  t.cfg = *cfg;
  t.row = t.hg.row;
  t.slots = 1ull << cfg->table_log2;
  uint32_t k = 0;
  while (k < c->tenants.size() && c->tenants[k].used) ++k;
  if (k >= kMaxTenants) return fail(c, MCCFR_E_ARG, "too many tenants (4096)");
  const size_t e = t.elems();
  if (!mem) {
    HIPX(c, hipMalloc(&t.owned, e * 10));
    mem = t.owned;
  }
  t.hg.R = static_cast<float*>(mem);  // interleaved rows: [slot][regrets | strategy sums]
  t.hg.S = t.hg.R + t.row;
  t.hg.sig = reinterpret_cast<uint16_t*>(t.hg.R + 2 * e);
  launch_zero(c->stream, mem, e * 10);
  launch_set_tenant(c->stream, c->d_tenants + k, t.hg);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    if (t.owned) (void)hipFree(t.owned);
    return fail(c, MCCFR_E_HIP, std::string("tenant set-up: ") + hipGetErrorString(err));
  }
  t.used = true;
  if (k == c->tenants.size()) c->tenants.push_back(t);
  else c->tenants[k] = t;
  *id = k;
  return MCCFR_OK;
}

int tenant_destroy(mccfr_ctx* c, uint32_t id) {
  Tenant* t = get_tenant(c, id);
  if (!t) return fail(c, MCCFR_E_ARG, "no such tenant");
  if (t->owned) {
    HIPX(c, hipStreamSynchronize(c->stream));
    HIPX(c, hipFree(t->owned));
  }
  *t = Tenant();
  drop_graph(c);  // the graph may hold launches of this tenant id
  return MCCFR_OK;
}

int issue_launch(mccfr_ctx* c, const HostSeg* segs, uint32_t n, mccfr_trace* host_trace, hipEvent_t done,
                 const DevScratch* scratch) {
  if (scratch && host_trace) return fail(c, MCCFR_E_ARG, "no trace for a launch on outside scratch");
  Prepared p;
  int rc = prepare(c, segs, n, scratch ? *scratch : c->sc, p);
  if (rc) return rc;
  if (p.lc.es) {
    if (scratch) return fail(c, MCCFR_E_ARG, "external sampling runs on the context's scratch only");
    if ((rc = ensure_es_stack(c, p.row, p.lc.lanes))) return rc;
  }
  const DevScratch& sc = scratch ? *scratch : c->sc;
  for (uint32_t i = 0; i < n; ++i)
    if ((rc = rm_sweep_if_dirty(c, c->tenants[segs[i].tenant]))) return rc;
  if (host_trace && !c->dtrace) HIPX(c, hipMalloc(&c->dtrace, (size_t)c->sc.cap_lanes * sizeof(mccfr_trace)));
  if ((rc = stage_segments(c, p.segs))) return rc;
  if ((rc = record_event(c, 0))) return rc;
  launch_walk(c->stream, p.row, p.lc, c->d_segs, c->d_tenants, deal_inputs(c), sc, c->shape,
              host_trace ? c->dtrace : nullptr, timing_event(c, 1), timing_event(c, 2));
  launch_reduce(c->stream, p.row, p.lc, c->d_segs, c->d_tenants, sc, p.sort_bits);
  if ((rc = record_event(c, 3))) return rc;
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    Tenant& t = c->tenants[segs[i].tenant];
    if (t.clamp_pending) {  // the reference clamps every slot after every launch (cfr_gpu.cu:1018-1022)
      launch_clamp(c->stream, t.row, t.hg.R, t.slots);
      t.clamp_pending = false;
    }
    t.traversals += segs[i].n;
    ++t.launches;
    total += segs[i].n;
  }
  HIPX(c, hipGetLastError());
  if (done) HIPX(c, hipEventRecord(done, c->stream));
  if (c->timing) {
    ++c->ev_next;
    c->ev_used = std::min(c->ev_used + 1, kTimingRing);
  }
  c->traversals += total;
  ++c->launches;
  if (host_trace) {
    size_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
      HIPX(c, hipMemcpyAsync(host_trace + off, c->dtrace + p.segs[i].lane0, (size_t)segs[i].n * sizeof(mccfr_trace),
                             hipMemcpyDeviceToHost, c->stream));
      off += segs[i].n;
    }
    HIPX(c, hipStreamSynchronize(c->stream));
  }
  if (p.lc.es && !done) return check_es(c);  // (the server checks its own launches)
  return MCCFR_OK;
}

int tenant_normalized(mccfr_ctx* c, uint32_t id, float* out, size_t elems, int layout) {
  Tenant* t = get_tenant(c, id);
  if (!t || !out) return fail(c, MCCFR_E_ARG, "no such tenant or null output");
  const size_t want = t->elems();
  if (elems != want) return fail(c, MCCFR_E_ARG, "elems must equal slots * row_actions");
  if (layout != 0 && layout != 1) return fail(c, MCCFR_E_ARG, "layout must be 0 (rows) or 1 (planes)");
  float* d = nullptr;
  HIPX(c, hipMalloc(&d, want * sizeof(float)));
  launch_normalize(c->stream, t->row, t->hg.S, d, t->slots);
  std::vector<float> tmp(layout ? want : 0);
  hipError_t e = hipMemcpyAsync(layout ? tmp.data() : out, d, want * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, MCCFR_E_HIP, hipGetErrorString(e));
  if (layout) to_soa(tmp.data(), out, t->slots, t->row);
  return MCCFR_OK;
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
  {
    DevGame probe;
    const char* why = nullptr;
    if (!make_dev_game(*cfg, nullptr, probe, &why)) return fail(nullptr, MCCFR_E_ARG, why);
  }
  auto* c = new (std::nothrow) mccfr_ctx();
  if (!c) return fail(nullptr, MCCFR_E_NOMEM, "host allocation failed");
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
  CHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  CHK(alloc_scratch(c->sc, round_up(max_batch, kSegAlign)));
  CHK(hipMalloc(&c->d_tenants, (size_t)kMaxTenants * sizeof(DevGame)));
  CHK(hipMalloc(&c->d_segs, (size_t)kMaxSegs * sizeof(DevSeg)));
  CHK(hipHostMalloc(&c->h_stage, (size_t)kStageRing * kMaxSegs * sizeof(DevSeg)));
  for (auto& e : c->stage_ev) CHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  CHK(hipMalloc(&c->lut, 2704));
  CHK(hipMalloc(&c->buckets.flop5, kFlopSets));
  CHK(hipMalloc(&c->buckets.turn6, kTurnSets));
  uint8_t lut[2704];
  build_preflop_lut(lut);
  CHK(hipMemcpy(c->lut, lut, sizeof lut, hipMemcpyHostToDevice));
  std::vector<uint16_t> nf, fl, off;
  build_arith_tables(nf, fl, off);
  CHK(hipMalloc(&c->arith.nonflush, nf.size() * 2));
  CHK(hipMalloc(&c->arith.flush, fl.size() * 2));
  CHK(hipMalloc(&c->arith.offsets, off.size() * 2));
  CHK(hipMemcpy(c->arith.nonflush, nf.data(), nf.size() * 2, hipMemcpyHostToDevice));
  CHK(hipMemcpy(c->arith.flush, fl.data(), fl.size() * 2, hipMemcpyHostToDevice));
  CHK(hipMemcpy(c->arith.offsets, off.data(), off.size() * 2, hipMemcpyHostToDevice));
#undef CHK
  uint32_t id = 0;
  int rc = tenant_create(c, cfg, nullptr, nullptr, &id);  // tenant 0: the context's own game
  if (rc == MCCFR_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = MCCFR_E_HIP;
  if (rc) return bail(rc);
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
  const char* why = "";
  if (table_check(hr, words, &why) != MCCFR_OK) return fail(c, MCCFR_E_TABLE, std::string("incompatible hand table: ") + why);
  if (!c->hr) HIPX(c, hipMalloc(&c->hr, words * sizeof(int32_t)));
  HIPX(c, hipStreamSynchronize(c->stream));
  HIPX(c, hipMemcpy(c->hr, hr, words * sizeof(int32_t), hipMemcpyHostToDevice));
  launch_build_buckets(c->stream, c->hr, c->buckets);  // flop/turn buckets of every 5-/6-card set
  c->eval_bk_ready = false;                            // the harness's buckets follow the table
  HIPX(c, hipGetLastError());
  HIPX(c, hipStreamSynchronize(c->stream));
  return MCCFR_OK;
}

int mccfr_load_hand_table(mccfr_ctx* c, const char* path) {
  if (!c || !path) return fail(c, MCCFR_E_ARG, "null argument");
  std::vector<int32_t> v;
  std::string msg;
  const int rc = read_table_file(path, v, msg);
  if (rc != MCCFR_OK) return fail(c, rc, msg);
  return mccfr_set_hand_table(c, v.data(), v.size());
}

int mccfr_validate_hand_table(const int32_t* hr, size_t words) {
  if (!hr) return fail(nullptr, MCCFR_E_ARG, "null argument");
  const char* why = "";
  const int rc = table_check(hr, words, &why);
  return rc == MCCFR_OK ? rc : fail(nullptr, rc, std::string("incompatible hand table: ") + why);
}

int mccfr_validate_hand_table_file(const char* path) {
  if (!path) return fail(nullptr, MCCFR_E_ARG, "null argument");
  std::vector<int32_t> v;
  std::string msg;
  const int rc = read_table_file(path, v, msg);
  return rc == MCCFR_OK ? rc : fail(nullptr, rc, msg);
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

int mccfr_run_launch(mccfr_ctx* c, int update_player, int64_t t_weight, uint64_t first_tid, uint32_t n,
                     mccfr_trace* trace) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (n == 0 || n > c->max_batch) return fail(c, MCCFR_E_ARG, "n must be 1..max_batch (see mccfr_set_max_batch)");
  HostSeg s{};
  s.tenant = 0;
  s.n = n;
  s.update_player = update_player;
  s.t_weight = t_weight;
  s.first_tid = first_tid;
  s.iteration = c->tenants[0].next_iter;
  s.upd_idx = 0;
  return issue_launch(c, &s, 1, trace, nullptr);
}

// The static arena's device address on the context's device, taken with
// hipGetSymbolAddress on the first call there (that device made current for
// the call, the caller's current device restored after it).
static int noalloc_scratch(mccfr_ctx* c, char** out) {
  if (c->device < 0 || c->device >= kMaxNoAllocDevices)
    return fail(c, MCCFR_E_STATE, "device index beyond the allocation-free launch's table");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[c->device]) {
    int cur = 0;
    HIPX(c, hipGetDevice(&cur));
    if (cur != c->device) HIPX(c, hipSetDevice(c->device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess && ptr) {
      // the emission block counts start at zero (the walk adds to them, the
      // emission clears what it read): a fill kernel on the context's stream
      // ahead of its first launch here
      static_assert((kNoAllocLanes / kEmitLanes * 4) % 16 == 0, "whole 16-byte chunks");
      DevScratch z;
      carve_scratch(z, static_cast<char*>(ptr), kNoAllocLanes);
      launch_zero(c->stream, z.blk_cnt, kNoAllocLanes / kEmitLanes * 4);
    }
    if (cur != c->device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !ptr)
      return fail(c, MCCFR_E_HIP, std::string("static launch scratch: ") + hipGetErrorString(e));
    g_noalloc_ptr[c->device] = static_cast<char*>(ptr);
  }
  *out = g_noalloc_ptr[c->device];
  return MCCFR_OK;
}

int mccfr_run_launch_noalloc(mccfr_ctx* c, int update_player, int64_t t_weight, uint64_t first_tid, uint32_t n) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (n == 0 || n > kNoAllocLanes)
    return fail(c, MCCFR_E_ARG, "n must be 1..16384 for mccfr_run_launch_noalloc (its static scratch)");
  char* base = nullptr;
  const int rc = noalloc_scratch(c, &base);
  if (rc) return rc;
  DevScratch sc;
  carve_scratch(sc, base, kNoAllocLanes);  // host arithmetic on the device address
  HostSeg s{};
  s.tenant = 0;
  s.n = n;
  s.update_player = update_player;
  s.t_weight = t_weight;
  s.first_tid = first_tid;
  s.iteration = c->tenants[0].next_iter;
  s.upd_idx = 0;
  return issue_launch(c, &s, 1, nullptr, nullptr, &sc);
}

int mccfr_run_segments(mccfr_ctx* c, const mccfr_segment* segs, uint32_t n_segs, mccfr_trace* trace) {
  if (!c || (!segs && n_segs)) return fail(c, MCCFR_E_ARG, "null argument");
  std::vector<HostSeg> hs(n_segs);
  for (uint32_t i = 0; i < n_segs; ++i) {
    hs[i] = HostSeg{};
    hs[i].tenant = segs[i].tenant;
    hs[i].n = segs[i].n;
    hs[i].update_player = segs[i].update_player;
    hs[i].t_weight = segs[i].t_weight;
    hs[i].first_tid = segs[i].first_tid;
    const Tenant* t = get_tenant(c, segs[i].tenant);
    hs[i].iteration = t ? t->next_iter : 1;
  }
  return issue_launch(c, hs.data(), n_segs, trace, nullptr);
}

int mccfr_train_tenants(mccfr_ctx* c, const uint32_t* tenants, const uint32_t* counts, uint32_t n, int64_t launches) {
  if (!c || !tenants || !counts) return fail(c, MCCFR_E_ARG, "null argument");
  if (launches <= 0) return MCCFR_OK;
  std::vector<HostSeg> hs(n);
  for (uint32_t i = 0; i < n; ++i) {
    const Tenant* t = get_tenant(c, tenants[i]);
    if (!t) return fail(c, MCCFR_E_ARG, "no such tenant");
    hs[i] = next_of(tenants[i], *t, counts[i]);
  }
  int rc;
  bool clamp = false;
  for (uint32_t i = 0; i < n; ++i) clamp = clamp || c->tenants[tenants[i]].clamp_pending;
  if (clamp) {  // the launch that follows an upload also clamps the whole table: issue it eagerly
    if ((rc = issue_launch(c, hs.data(), n, nullptr, nullptr))) return rc;
    for (uint32_t i = 0; i < n; ++i) advance(c->tenants[tenants[i]], hs[i]);
    if (--launches == 0) goto done;
  }
  {
    Prepared p;
    if ((rc = prepare(c, hs.data(), n, c->sc, p))) return rc;
    if (p.lc.es && (rc = ensure_es_stack(c, p.row, p.lc.lanes))) return rc;
    for (uint32_t i = 0; i < n; ++i)
      if ((rc = rm_sweep_if_dirty(c, c->tenants[tenants[i]]))) return rc;
    std::vector<uint64_t> key = {(uint64_t)p.row,          (uint64_t)p.lc.lanes, (uint64_t)p.lc.n_segs,
                                 (uint64_t)p.lc.slot_bits, (uint64_t)p.key_bits, (uint64_t)c->shape.wg,
                                 (uint64_t)c->shape.tpl,   (uint64_t)c->sc.cap_lanes};
    for (const DevSeg& d : p.segs) key.push_back(((uint64_t)d.tenant << 32) | d.n);
    if ((rc = stage_segments(c, p.segs))) return rc;
    if (!c->gexec || c->graph_key != key) {
      drop_graph(c);
      // one batched launch: deal + walk -> records -> sort -> table update, run
      // discovery moving the segment table on to the next launch
      hipGraph_t g = nullptr;
      HIPX(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      launch_walk(c->stream, p.row, p.lc, c->d_segs, c->d_tenants, deal_inputs(c), c->sc, c->shape, nullptr);
      launch_reduce(c->stream, p.row, p.lc, c->d_segs, c->d_tenants, c->sc, p.sort_bits, c->d_segs);
      HIPX(c, hipStreamEndCapture(c->stream, &g));
      hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        c->gexec = nullptr;
        return fail(c, MCCFR_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
      c->graph_key = key;
    }
    for (int64_t L = 0; L < launches; ++L) HIPX(c, hipGraphLaunch(c->gexec, c->stream));
    uint64_t per = 0;
    for (uint32_t i = 0; i < n; ++i) per += counts[i];
    c->traversals += per * (uint64_t)launches;
    c->launches += (uint64_t)launches;
    for (int64_t L = 0; L < launches; ++L)
      for (uint32_t i = 0; i < n; ++i) {
        Tenant& t = c->tenants[tenants[i]];
        t.traversals += counts[i];
        ++t.launches;
        advance(t, hs[i]);
      }
  }
done:
  for (uint32_t i = 0; i < n; ++i) {
    Tenant& t = c->tenants[tenants[i]];
    t.next_iter = hs[i].iteration;
    t.upd_idx = hs[i].upd_idx;
  }
  HIPX(c, hipStreamSynchronize(c->stream));
  return check_es(c);
}

int mccfr_train(mccfr_ctx* c, int64_t iters, uint32_t batch) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (!c->hr) return fail(c, MCCFR_E_STATE, "no hand table loaded");
  if (iters <= 0) return MCCFR_OK;
  if (batch == 0 || batch > c->max_batch) return fail(c, MCCFR_E_ARG, "batch must be 1..max_batch");
  Tenant& t = c->tenants[0];
  t.upd_idx = 0;  // for iter = 1..iters, for p = 0..N-1 (cfr_gpu.cu:1286-1322)
  const uint32_t id = 0;
  return mccfr_train_tenants(c, &id, &batch, 1, iters * t.hg.n_upd);
}

int mccfr_set_max_batch(mccfr_ctx* c, uint32_t max_batch) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (max_batch == 0 || max_batch > (1u << 26)) return fail(c, MCCFR_E_ARG, "max_batch must be 1..2^26");
  HIPX(c, hipStreamSynchronize(c->stream));
  drop_graph(c);  // the graph holds the old scratch pointers
  const uint32_t cap = round_up(max_batch, kSegAlign);
  if (cap != c->sc.cap_lanes) {
    free_scratch(c->sc);
    const hipError_t e = alloc_scratch(c->sc, cap);
    c->sc.es_stack = c->es_stack;
    c->sc.es_rd = c->es_rd;  // regrown for the new capacity by the next external-sampling launch
    if (e != hipSuccess) {
      free_scratch(c->sc);
      c->max_batch = 0;
      return fail(c, MCCFR_E_NOMEM, std::string("launch scratch: ") + hipGetErrorString(e));
    }
  }
  if (c->dtrace) {
    (void)hipFree(c->dtrace);
    c->dtrace = nullptr;
  }
  c->max_batch = max_batch;
  return MCCFR_OK;
}

static constexpr uint64_t kCheckpointMagic = 0x43465247505500ull;  // cfr_gpu.cu:1480

int mccfr_save_checkpoint(mccfr_ctx* c, const char* path) {
  if (!c || !path) return fail(c, MCCFR_E_ARG, "null argument");
  const Tenant& t = c->tenants[0];
  if (t.slots > 0xFFFFFFFFull) return fail(c, MCCFR_E_ARG, "table too large for the .ckpt header");
  const size_t n = t.elems();
  std::vector<float> R(n), S(n);
  int rc = mccfr_download_tables(c, R.data(), S.data(), nullptr, n, 1);
  if (rc) return rc;
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(c, MCCFR_E_IO, std::string("cannot open ") + path);
  const uint64_t magic = kCheckpointMagic;
  const uint32_t ts = (uint32_t)t.slots, na = (uint32_t)t.row;
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
  const Tenant& t = c->tenants[0];
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(c, MCCFR_E_IO, std::string("cannot open ") + path);
  uint64_t magic = 0;
  uint32_t ts = 0, na = 0;
  f.read(reinterpret_cast<char*>(&magic), 8);
  f.read(reinterpret_cast<char*>(&ts), 4);
  f.read(reinterpret_cast<char*>(&na), 4);
  if (!f) return fail(c, MCCFR_E_IO, std::string("cannot read header of ") + path);
  if (magic != kCheckpointMagic) return fail(c, MCCFR_E_FORMAT, std::string(path) + " is not a checkpoint");
  if ((uint64_t)ts != t.slots || na != (uint32_t)t.row)
    return fail(c, MCCFR_E_FORMAT, std::string(path) + " was saved with table_size=" + std::to_string(ts) +
                                       " num_actions=" + std::to_string(na) + ", this context has " +
                                       std::to_string(t.slots) + " and " + std::to_string(t.row));
  const size_t n = (size_t)ts * na;
  std::vector<float> R(n), S(n);
  f.read(reinterpret_cast<char*>(R.data()), (std::streamsize)(n * sizeof(float)));
  f.read(reinterpret_cast<char*>(S.data()), (std::streamsize)(n * sizeof(float)));
  if (!f) return fail(c, MCCFR_E_IO, std::string("truncated checkpoint ") + path);
  return mccfr_upload_tables(c, R.data(), S.data(), n, 1);
}

int mccfr_set_launch_shape(mccfr_ctx* c, uint32_t walk_wg, uint32_t tpl) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (walk_wg != 64 && walk_wg != 128 && walk_wg != 256 && walk_wg != 512 && walk_wg != 1024)
    return fail(c, MCCFR_E_ARG, "walk workgroup must be 64, 128, 256, 512 or 1024");
  if (tpl != 0 && tpl != 1 && tpl != 2 && tpl != 4)
    return fail(c, MCCFR_E_ARG, "traversals per lane must be 1, 2 or 4 (0: persistent)");
  if (c->shape.wg != walk_wg || c->shape.tpl != tpl) drop_graph(c);
  c->shape.wg = walk_wg;
  c->shape.tpl = tpl;
  return MCCFR_OK;
}

int mccfr_get_launch_shape(const mccfr_ctx* c, uint32_t* walk_wg, uint32_t* tpl) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (walk_wg) *walk_wg = c->shape.wg;
  if (tpl) *tpl = c->shape.tpl;
  return MCCFR_OK;
}

int mccfr_autotune(mccfr_ctx* c, const mccfr_segment* segs, uint32_t n_segs, int reps, uint32_t* walk_wg,
                   uint32_t* tpl, float* ms_out) {
  if (!c || !segs || reps < 1) return fail(c, MCCFR_E_ARG, "bad autotune arguments");
  std::vector<HostSeg> hs(n_segs);
  for (uint32_t i = 0; i < n_segs; ++i) {
    hs[i] = HostSeg{};
    hs[i].tenant = segs[i].tenant;
    hs[i].n = segs[i].n;
    hs[i].update_player = segs[i].update_player;
    hs[i].t_weight = segs[i].t_weight;
    hs[i].first_tid = segs[i].first_tid;
  }
  Prepared p;
  int rc = prepare(c, hs.data(), n_segs, c->sc, p);
  if (rc) return rc;
  if (p.lc.es) return fail(c, MCCFR_E_ARG, "the autotuner sweeps the outcome-sampling walk's shapes");
  for (uint32_t i = 0; i < n_segs; ++i)
    if ((rc = rm_sweep_if_dirty(c, c->tenants[segs[i].tenant]))) return rc;
  if ((rc = stage_segments(c, p.segs))) return rc;
  hipEvent_t e0, e1;
  HIPX(c, hipEventCreate(&e0));
  HIPX(c, hipEventCreate(&e1));
  const WalkShape keep = c->shape;
  WalkShape best = keep;
  float best_ms = 1e30f;
  const uint32_t wgs[5] = {64, 128, 256, 512, 1024};
  const uint32_t tpls[4] = {1, 2, 4, 0};
  hipError_t err = hipSuccess;
  for (uint32_t wg : wgs)
    for (uint32_t tp : tpls) {
      if (tp == 0 && p.lc.n_segs > 1) continue;  // persistent waves serve one-tenant launches only
      WalkShape sh;
      sh.wg = wg;
      sh.tpl = tp;
      launch_walk(c->stream, p.row, p.lc, c->d_segs, c->d_tenants, deal_inputs(c), c->sc, sh, nullptr);  // warm-up
      if (err == hipSuccess) err = hipEventRecord(e0, c->stream);
      for (int r = 0; r < reps; ++r)
        launch_walk(c->stream, p.row, p.lc, c->d_segs, c->d_tenants, deal_inputs(c), c->sc, sh, nullptr);
      if (err == hipSuccess) err = hipEventRecord(e1, c->stream);
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      float ms = 0.f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
      ms /= (float)reps;
      if (err == hipSuccess && ms < best_ms) {
        best_ms = ms;
        best = sh;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err != hipSuccess) return fail(c, MCCFR_E_HIP, std::string("autotune: ") + hipGetErrorString(err));
  if (best.wg != keep.wg || best.tpl != keep.tpl) drop_graph(c);
  c->shape = best;
  if (walk_wg) *walk_wg = best.wg;
  if (tpl) *tpl = best.tpl;
  if (ms_out) *ms_out = best_ms;
  return MCCFR_OK;
}

int mccfr_randomize_tables(mccfr_ctx* c, uint64_t seed) { return mccfr_tenant_randomize(c, 0, seed); }

int mccfr_tenant_randomize(mccfr_ctx* c, uint32_t id, uint64_t seed) {
  Tenant* t = get_tenant(c, id);
  if (!t) return fail(c, MCCFR_E_ARG, "no such tenant");
  launch_randomize(c->stream, t->row, t->cfg.num_actions, t->hg.R, t->hg.S, t->slots, seed);
  HIPX(c, hipGetLastError());
  t->sigma_dirty = true;
  return rm_sweep_if_dirty(c, *t);
}

int mccfr_sync(mccfr_ctx* c) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  HIPX(c, hipStreamSynchronize(c->stream));
  return MCCFR_OK;
}

int mccfr_table_shape(const mccfr_ctx* c, uint64_t* slots, uint32_t* row) {
  return mccfr_tenant_table_shape(c, 0, slots, row);
}

int mccfr_tenant_table_shape(const mccfr_ctx* c, uint32_t id, uint64_t* slots, uint32_t* row) {
  if (!c || id >= c->tenants.size() || !c->tenants[id].used) return fail(nullptr, MCCFR_E_ARG, "no such tenant");
  const Tenant& t = c->tenants[id];
  if (slots) *slots = t.slots;
  if (row) *row = (uint32_t)t.row;
  return MCCFR_OK;
}

// One plane (regrets or strategy sums) of a tenant between its interleaved
// device rows and dense host rows [slot][row], through a device staging buffer
// of at most 64 MB.  to_host: device -> host; else host -> device (the host
// buffer is only read).
static int plane_transfer(mccfr_ctx* c, const Tenant& t, float* plane, float* host, bool to_host) {
  const uint64_t chunk = std::min<uint64_t>(t.slots, (64ull << 20) / ((uint64_t)t.row * sizeof(float)));
  float* stage = nullptr;
  HIPX(c, hipMalloc(&stage, chunk * t.row * sizeof(float)));
  hipError_t e = hipSuccess;
  for (uint64_t s0 = 0; s0 < t.slots && e == hipSuccess; s0 += chunk) {
    const uint64_t n = std::min<uint64_t>(chunk, t.slots - s0);
    const size_t bytes = n * t.row * sizeof(float);
    if (to_host) {
      launch_plane(c->stream, t.row, plane, stage, s0, n, true);
      e = hipMemcpyAsync(host + s0 * t.row, stage, bytes, hipMemcpyDeviceToHost, c->stream);
    } else {
      e = hipMemcpyAsync(stage, host + s0 * t.row, bytes, hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess) launch_plane(c->stream, t.row, plane, stage, s0, n, false);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // the stage is reused by the next chunk
  }
  (void)hipFree(stage);
  if (e != hipSuccess) return fail(c, MCCFR_E_HIP, std::string("table transfer: ") + hipGetErrorString(e));
  return MCCFR_OK;
}

int mccfr_download_tables(mccfr_ctx* c, float* R, float* S, uint16_t* sig, size_t elems, int layout) {
  return mccfr_tenant_download(c, 0, R, S, sig, elems, layout);
}

int mccfr_tenant_download(mccfr_ctx* c, uint32_t id, float* R, float* S, uint16_t* sig, size_t elems, int layout) {
  Tenant* t = get_tenant(c, id);
  if (!t) return fail(c, MCCFR_E_ARG, "no such tenant");
  const size_t want = t->elems();
  if (elems != want) return fail(c, MCCFR_E_ARG, "elems must equal slots * row_actions");
  if (layout != 0 && layout != 1) return fail(c, MCCFR_E_ARG, "layout must be 0 (rows) or 1 (planes)");
  int rc = rm_sweep_if_dirty(c, *t);
  if (rc) return rc;
  HIPX(c, hipStreamSynchronize(c->stream));
  auto get = [&](float* host, float* plane) -> int {
    if (!host) return MCCFR_OK;
    if (layout == 0) return plane_transfer(c, *t, plane, host, true);
    std::vector<float> tmp(want);
    if ((rc = plane_transfer(c, *t, plane, tmp.data(), true))) return rc;
    to_soa(tmp.data(), host, t->slots, t->row);
    return MCCFR_OK;
  };
  if ((rc = get(R, t->hg.R))) return rc;
  if ((rc = get(S, t->hg.S))) return rc;
  if (sig) {
    if (layout == 0) {
      HIPX(c, hipMemcpy(sig, t->hg.sig, want * sizeof(uint16_t), hipMemcpyDeviceToHost));
    } else {
      std::vector<uint16_t> tmp(want);
      HIPX(c, hipMemcpy(tmp.data(), t->hg.sig, want * sizeof(uint16_t), hipMemcpyDeviceToHost));
      to_soa(tmp.data(), sig, t->slots, t->row);
    }
  }
  return MCCFR_OK;
}

int mccfr_upload_tables(mccfr_ctx* c, const float* R, const float* S, size_t elems, int layout) {
  return mccfr_tenant_upload(c, 0, R, S, elems, layout);
}

int mccfr_tenant_upload(mccfr_ctx* c, uint32_t id, const float* R, const float* S, size_t elems, int layout) {
  Tenant* t = get_tenant(c, id);
  if (!t) return fail(c, MCCFR_E_ARG, "no such tenant");
  const size_t want = t->elems();
  if (elems != want) return fail(c, MCCFR_E_ARG, "elems must equal slots * row_actions");
  if (layout != 0 && layout != 1) return fail(c, MCCFR_E_ARG, "layout must be 0 (rows) or 1 (planes)");
  HIPX(c, hipStreamSynchronize(c->stream));
  auto put = [&](const float* host, float* plane) -> int {
    if (!host) return MCCFR_OK;
    if (layout == 0) return plane_transfer(c, *t, plane, const_cast<float*>(host), false);
    std::vector<float> tmp(want);
    to_aos(host, tmp.data(), t->slots, t->row);
    return plane_transfer(c, *t, plane, tmp.data(), false);
  };
  int rc;
  if ((rc = put(R, t->hg.R))) return rc;
  if ((rc = put(S, t->hg.S))) return rc;
  t->sigma_dirty = true;  // the next launch regret-matches the whole table first (cfr_gpu.cu:1290)
  if (R && t->cfg.use_cfr_plus) t->clamp_pending = true;
  return MCCFR_OK;
}

static constexpr uint64_t kStrategyMagicV1 = 0x53545241540000ull;  // strategy.cpp:10
static constexpr uint64_t kStrategyMagicV2 = 0x53545241540100ull;  // strategy.cpp:11

int mccfr_write_strategy(const char* path, uint32_t table_size, uint32_t num_actions, uint32_t entries,
                         const uint32_t* keys, const float* probs) {
  if (!path || (entries && (!keys || !probs)) || num_actions == 0 || num_actions > MCCFR_MAX_ACTIONS)
    return fail(nullptr, MCCFR_E_ARG, "bad strategy write arguments");
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(nullptr, MCCFR_E_IO, std::string("cannot open ") + path);
  const uint64_t magic = kStrategyMagicV2;
  f.write(reinterpret_cast<const char*>(&magic), 8);
  f.write(reinterpret_cast<const char*>(&table_size), 4);
  f.write(reinterpret_cast<const char*>(&num_actions), 4);
  f.write(reinterpret_cast<const char*>(&entries), 4);
  for (uint32_t i = 0; i < entries; ++i) {
    f.write(reinterpret_cast<const char*>(&keys[i]), 4);
    f.write(reinterpret_cast<const char*>(probs + (size_t)i * num_actions), (std::streamsize)(num_actions * 4));
  }
  if (!f) return fail(nullptr, MCCFR_E_IO, std::string("write failed: ") + path);
  return MCCFR_OK;
}

int mccfr_save_strategy(mccfr_ctx* c, uint32_t id, const char* path) {
  if (!c || !path) return fail(c, MCCFR_E_ARG, "null argument");
  Tenant* t = get_tenant(c, id);
  if (!t) return fail(c, MCCFR_E_ARG, "no such tenant");
  if (t->slots > 0xFFFFFFFFull) return fail(c, MCCFR_E_ARG, "table too large for the .bin header");
  const size_t n = t->elems();
  std::vector<float> P(n);
  int rc = tenant_normalized(c, id, P.data(), n, 0);
  if (rc) return rc;
  // get_strategy's row selection (cfr_gpu.cu:1383-1412), ascending slot order
  std::vector<uint32_t> keys;
  size_t kept = 0;
  for (uint64_t s = 0; s < t->slots; ++s) {
    const float* row = &P[s * t->row];
    float total = 0.f;
    for (int a = 0; a < t->row; ++a) total += row[a];
    if (total < 1e-6f) continue;
    keys.push_back((uint32_t)s);
    if (kept != s) std::memmove(&P[kept * t->row], row, t->row * sizeof(float));
    ++kept;
  }
  rc = mccfr_write_strategy(path, (uint32_t)t->slots, (uint32_t)t->row, (uint32_t)keys.size(), keys.data(),
                            P.data());
  if (rc) c->err = mccfr_last_error(nullptr);
  return rc;
}

int mccfr_load_strategy(const char* path, uint32_t* table_size, uint32_t* num_actions, uint32_t* entries,
                        uint32_t* keys, float* probs, size_t capacity) {
  if (!path) return fail(nullptr, MCCFR_E_ARG, "null path");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(nullptr, MCCFR_E_IO, std::string("cannot open ") + path);
  uint64_t magic = 0;
  f.read(reinterpret_cast<char*>(&magic), 8);
  if (!f) return fail(nullptr, MCCFR_E_IO, std::string("cannot read ") + path);
  if (magic == kCheckpointMagic)
    return fail(nullptr, MCCFR_E_FORMAT, std::string(path) + " is a training checkpoint, not a strategy file");
  if (magic == kStrategyMagicV1)
    return fail(nullptr, MCCFR_E_FORMAT, std::string(path) + " is an older (v1) strategy export");
  if (magic != kStrategyMagicV2)
    return fail(nullptr, MCCFR_E_FORMAT, std::string(path) + " does not have a strategy file header");
  uint32_t ts = 0, na = 0, n = 0;
  f.read(reinterpret_cast<char*>(&ts), 4);
  f.read(reinterpret_cast<char*>(&na), 4);
  f.read(reinterpret_cast<char*>(&n), 4);
  if (!f || na == 0 || na > MCCFR_MAX_ACTIONS) return fail(nullptr, MCCFR_E_FORMAT, std::string("bad header in ") + path);
  if (table_size) *table_size = ts;
  if (num_actions) *num_actions = na;
  std::vector<float> row(na);
  const float uniform = 1.0f / (float)na;
  uint32_t kept = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t k = 0;
    f.read(reinterpret_cast<char*>(&k), 4);
    f.read(reinterpret_cast<char*>(row.data()), (std::streamsize)(na * sizeof(float)));
    if (!f) return fail(nullptr, MCCFR_E_FORMAT, std::string("truncated strategy file ") + path);
    bool placeholder = true;  // strategy.cpp:22-33
    for (uint32_t a = 0; a < na; ++a) placeholder = placeholder && std::fabs(row[a] - uniform) <= 1e-6f;
    if (placeholder) continue;
    if (keys) {
      if (kept >= capacity) return fail(nullptr, MCCFR_E_ARG, "capacity smaller than the file's entries");
      keys[kept] = k;
      if (probs) std::memcpy(probs + (size_t)kept * na, row.data(), na * sizeof(float));
    }
    ++kept;
  }
  if (entries) *entries = kept;
  return MCCFR_OK;
}

int mccfr_normalized_strategy(mccfr_ctx* c, float* out, size_t elems, int layout) {
  return tenant_normalized(c, 0, out, elems, layout);
}

int mccfr_tenant_normalized_strategy(mccfr_ctx* c, uint32_t id, float* out, size_t elems, int layout) {
  return tenant_normalized(c, id, out, elems, layout);
}

int mccfr_tenant_footprint(const mccfr_game_cfg* cfg, uint64_t* bytes) {
  if (!cfg || !bytes) return fail(nullptr, MCCFR_E_ARG, "null argument");
  DevGame g;
  const char* why = nullptr;
  if (!make_dev_game(*cfg, nullptr, g, &why)) return fail(nullptr, MCCFR_E_ARG, why);
  *bytes = tenant_bytes(*cfg);
  return MCCFR_OK;
}

int mccfr_tenant_create(mccfr_ctx* c, const mccfr_game_cfg* cfg, const mccfr_root_state* root, uint32_t* tenant) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  int rc = tenant_create(c, cfg, root, nullptr, tenant);
  if (rc) return rc;
  HIPX(c, hipStreamSynchronize(c->stream));
  return MCCFR_OK;
}

int mccfr_tenant_destroy(mccfr_ctx* c, uint32_t tenant) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (tenant == 0) return fail(c, MCCFR_E_ARG, "tenant 0 is the context's own game");
  return tenant_destroy(c, tenant);
}

int mccfr_tenant_counters(const mccfr_ctx* c, uint32_t id, uint64_t* traversals, uint64_t* launches,
                          int64_t* next_iteration, int32_t* next_player) {
  if (!c || id >= c->tenants.size() || !c->tenants[id].used) return fail(nullptr, MCCFR_E_ARG, "no such tenant");
  const Tenant& t = c->tenants[id];
  if (traversals) *traversals = t.traversals;
  if (launches) *launches = t.launches;
  if (next_iteration) *next_iteration = t.next_iter;
  if (next_player) *next_player = t.hg.upd_seats[t.upd_idx];
  return MCCFR_OK;
}

int mccfr_counters(const mccfr_ctx* c, uint64_t* traversals, uint64_t* launches, int64_t* next_iteration) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (traversals) *traversals = c->traversals;
  if (launches) *launches = c->launches;
  if (next_iteration) *next_iteration = c->tenants[0].next_iter;
  return MCCFR_OK;
}

int mccfr_set_timing(mccfr_ctx* c, int enabled) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (enabled && c->ev.empty()) {
    c->ev.resize((size_t)kTimingRing * 4, nullptr);
    for (auto& e : c->ev) HIPX(c, hipEventCreate(&e));
  }
  c->timing = enabled != 0;
  c->ev_next = 0;
  c->ev_used = 0;
  return MCCFR_OK;
}

int mccfr_kernel_times(mccfr_ctx* c, float* deal_ms, float* walk_ms, float* upd_ms, uint32_t n) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  if (n > c->ev_used) return fail(c, MCCFR_E_ARG, "fewer timed launches recorded than requested");
  HIPX(c, hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t k = (c->ev_next - n + i) % kTimingRing;
    float a = 0.f, b = 0.f, d = 0.f;
    HIPX(c, hipEventElapsedTime(&a, c->ev[k * 4 + 0], c->ev[k * 4 + 1]));
    HIPX(c, hipEventElapsedTime(&b, c->ev[k * 4 + 1], c->ev[k * 4 + 2]));
    HIPX(c, hipEventElapsedTime(&d, c->ev[k * 4 + 2], c->ev[k * 4 + 3]));
    if (deal_ms) deal_ms[i] = a;
    if (walk_ms) walk_ms[i] = b;
    if (upd_ms) upd_ms[i] = d;
  }
  return MCCFR_OK;
}

int mccfr_last_launch_stats(mccfr_ctx* c, uint32_t* records, uint32_t* long_runs, uint32_t* pieces) {
  if (!c) return fail(nullptr, MCCFR_E_ARG, "null context");
  LaunchCounters h{};
  HIPX(c, hipStreamSynchronize(c->stream));
  HIPX(c, hipMemcpy(&h, c->sc.cnt, sizeof h, hipMemcpyDeviceToHost));
  if (records) *records = h.n_rec;
  if (long_runs) *long_runs = h.n_long;
  if (pieces) *pieces = h.n_pieces;
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
  launch_zero(c->stream, a, bytes);
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

int mccfr_gather_probe(mccfr_ctx* c, size_t table_bytes, size_t reads, int reps, float* ms_per_pass) {
  if (!c || reads == 0 || reps < 1 || table_bytes < 16 || (table_bytes & (table_bytes - 1)))
    return fail(c, MCCFR_E_ARG, "bad gather probe arguments (table_bytes a power of two >= 16)");
  void *a = nullptr, *b = nullptr;
  HIPX(c, hipMalloc(&a, table_bytes));
  hipError_t e = hipMalloc(&b, reads * sizeof(uint32_t));
  if (e != hipSuccess) {
    (void)hipFree(a);
    return fail(c, MCCFR_E_NOMEM, "device allocation failed");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  launch_zero(c->stream, a, table_bytes);
  launch_gather(c->stream, (const uint4*)a, table_bytes / 16, (uint32_t*)b, reads);  // warm-up
  (void)hipEventRecord(e0, c->stream);
  for (int r = 0; r < reps; ++r) launch_gather(c->stream, (const uint4*)a, table_bytes / 16, (uint32_t*)b, reads);
  (void)hipEventRecord(e1, c->stream);
  e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(c, MCCFR_E_HIP, hipGetErrorString(e));
  if (ms_per_pass) *ms_per_pass = ms / reps;
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
