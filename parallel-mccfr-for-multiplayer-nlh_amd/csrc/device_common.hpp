// Device-side building blocks shared by the gfx950 kernels: the game
// descriptor, Philox streams, BF16 rounding, the info-set hash and the
// betting rules.  Every function is __host__ __device__ so the host runtime
// can use the same rule code for validation; the parity oracle is a separate
// CPU restatement under oracle/.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/mccfr/mccfr.h"

namespace mccfr {

constexpr int kMaxPlayers = 6;
constexpr int kMaxDepth = 20;  // MAX_TRAVERSE_DEPTH, cfr_gpu.cu:642
constexpr uint32_t kHistoryMask = 0x3FFFFFFFu;  // cfr_gpu.cu:553

// ---------------------------------------------------------------------------
// Tenant descriptor: one game configuration, the root its traversals start
// from, its tables and its update-player schedule.  The context's own game is
// tenant 0; a batched launch runs segments of several tenants.
// ---------------------------------------------------------------------------
struct DevGame {
  int32_t N;
  int32_t sb, bb;
  int32_t num_actions;
  int32_t row;          // padded actions per table row (8 or 16)
  int32_t max_raises;   // 0 = unlimited
  int32_t hist_bits;
  int32_t table_log2;
  int32_t use_cfr_plus;
  int32_t use_linear_cfr;
  int32_t sampling;      // MCCFR_SAMPLE_OUTCOME or MCCFR_SAMPLE_EXTERNAL
  uint32_t seed_lo, seed_hi;
  uint32_t tag;          // Philox counter word 3 (stream family)
  // stack[p]: chips of seat p at the start of the hand, so that invested =
  // stack[p] - chips behind (preflop root: the starting stack; postflop root:
  // chips behind at the root + chips invested before it).
  int32_t stack[kMaxPlayers];
  uint8_t kind[MCCFR_MAX_ACTIONS];
  uint8_t num[MCCFR_MAX_ACTIONS];
  uint8_t den[MCCFR_MAX_ACTIONS];
  uint32_t passive_order[MCCFR_MAX_ACTIONS];  // fallback preference list (menu indices)
  int32_t n_passive;
  uint32_t check_mask, call_mask, fold_mask, allin_mask, raise_mask;  // menu indices by kind

  // root of the tree (street 0 = the reference's preflop root)
  int32_t root_street;
  int32_t root_dealer;        // postflop roots
  uint32_t root_folded;       // seats out of the hand at the root
  uint32_t root_allin;        // seats with no chips behind at the root
  uint32_t root_ab;           // action-history bits at the root
  int32_t root_pot;
  int32_t root_stack[kMaxPlayers];  // chips behind at the root
  uint32_t board_lo, board_hi;      // fixed board cards as a 52-bit mask
  uint8_t board[5];
  uint8_t n_board;            // fixed board cards: 0, 3, 4 or 5
  uint8_t deck_size;          // 52 - n_board
  uint8_t n_deal;             // cards dealt per traversal: 2N + 5 - n_board
  uint32_t walk_k0;           // first draw index of the walk (after the deal draws)
  uint8_t upd_seats[kMaxPlayers];  // update-player schedule
  int32_t n_upd;

  // tables: slot s's regrets and strategy sums share one interleaved row,
  // R[s * 2 * row + a] and S[s * 2 * row + a] (S = R + row), so the update
  // reads and writes both with one access; the BF16 snapshot the walk gathers
  // is a dense [slot][row] table of its own
  float* R;                   // regrets
  float* S;                   // strategy sums
  uint16_t* sig;              // BF16 regret-matched strategy snapshot
};

// One tenant's share of a launch: lanes [lane0, lane0 + n) of the launch's
// lane space; lane0 is a multiple of kSegAlign so every wave's lanes belong
// to one segment.  The captured training graph advances it on the device.
struct DevSeg {
  uint32_t tenant;
  uint32_t lane0;
  uint32_t n;
  int32_t update_player;
  int64_t iteration;     // training-loop iteration (1-based)
  int64_t t_weight;      // strategy-sum weight: iteration under linear CFR, else 1
  uint64_t first_tid;    // traversal id of the segment's first lane
  uint32_t upd_idx;      // position of update_player in the tenant's schedule
  uint32_t pad;
};
constexpr uint32_t kSegAlign = 256;

// A segment's next launch in the training loop (cfr_gpu.cu:1303-1318): the
// next traversal ids, the next seat of the tenant's update schedule (the first
// seat of the next iteration after the last) and its strategy-sum weight.
__host__ __device__ inline void advance_segment(DevSeg& s, const DevGame& G) {
  s.first_tid += s.n;
  if (++s.upd_idx >= (uint32_t)G.n_upd) {
    s.upd_idx = 0;
    s.iteration += 1;
  }
  s.update_player = G.upd_seats[s.upd_idx];
  s.t_weight = G.use_linear_cfr ? s.iteration : 1;
}

// Per-launch constants passed by value to the kernels.
struct LaunchCfg {
  uint32_t n_segs;
  uint32_t lanes;        // lane space of the launch (sum of aligned segment sizes)
  uint32_t slot_bits;    // record key = segment << slot_bits | slot
  uint32_t cap_lanes;    // lane capacity of the scratch lists
  uint32_t es;           // 1: the launch's tenants use external sampling (k_walk_es)
  uint32_t null_key;     // external sampling: key of a null record, above every real key
};

// Counters the kernels of one launch write (reset by the launch itself).
struct LaunchCounters {
  uint32_t n_rec;        // update-player records emitted
  uint32_t n_cand;       // runs whose end the per-wave pass could not see
  uint32_t n_long;       // runs handed to the long-run reduction
  uint32_t n_pieces;     // pieces of the long runs (with n_long: one 64-bit word, bumped together)
  uint32_t n_runs;       // short runs listed for the row-parallel reduction
  uint32_t n_es;         // external sampling: records reserved by the walk, in whole chunks
  uint32_t es_overflow;  // sticky: 1 records beyond capacity, 2 stack beyond its frames (host clears)
  uint32_t n_sort;       // records the radix passes sort (external sampling: the reservations)
  uint32_t n_es_valid;   // external sampling: real records among them (the rest are null records)
  uint32_t pad[2];
};
static_assert(offsetof(LaunchCounters, n_long) % 8 == 0 &&
                  offsetof(LaunchCounters, n_pieces) == offsetof(LaunchCounters, n_long) + 4,
              "n_long and n_pieces form one little-endian 64-bit word");

// What the table update needs of a segment, gathered once per launch.
struct SegCtx {
  float* R;
  float* S;
  uint16_t* sig;
  float tw;              // strategy-sum weight as FP32 (cfr_gpu.cu:828)
  int16_t num_actions;
  int16_t cfr_plus;
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Draw k of traversal tid
// is word k&3 of Philox(counter = {k>>2, tid_lo, tid_hi, tag}, key = {seed_lo, seed_hi}).
// ---------------------------------------------------------------------------
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Path code of the task under action a of an update-player node whose own
// task has code p (external sampling): one FNV-1a step, 26 bits, never 0.
__host__ __device__ inline uint32_t es_child_path(uint32_t p, int a) {
  const uint32_t h = ((p ^ (uint32_t)(a + 1)) * 16777619u) & 0x03FFFFFFu;
  return h ? h : 1u;
}

// External sampling splits a traversal's tree into tasks (the root's, and
// one per action of every update-player node), each with its own stream:
// `path` is the task's code (0 for the root and every outcome-sampling walk),
// held in the counter's first word above the block index (k < 256 on every
// walk).
struct DrawStream {
  uint32_t k0, k1, tid_lo, tid_hi, tag;
  uint32_t path;   // task path code (external sampling), 0 otherwise
  uint32_t k;      // next draw index
  uint32_t blk;    // block id held in w[] (0xFFFFFFFF = none)
  uint32_t w[4];
  __host__ __device__ void init(uint32_t seed_lo, uint32_t seed_hi, uint64_t tid, uint32_t tg) {
    k0 = seed_lo;
    k1 = seed_hi;
    tid_lo = (uint32_t)tid;
    tid_hi = (uint32_t)(tid >> 32);
    tag = tg;
    path = 0;
    k = 0;
    blk = 0xFFFFFFFFu;
  }
  // the stream of the task under action a of an update-player node of this task
  __host__ __device__ void enter_child(uint32_t node_path, int a) {
    path = es_child_path(node_path, a);
    k = 0;
    blk = 0xFFFFFFFFu;
  }
  __host__ __device__ uint32_t next() {
    const uint32_t b = k >> 2;
    if (b != blk) {
      philox4x32_10(b | (path << 6), tid_lo, tid_hi, tag, k0, k1, w);
      blk = b;
    }
    const uint32_t i = k & 3u;
    ++k;
    // select without dynamic register indexing
    return i == 0 ? w[0] : (i == 1 ? w[1] : (i == 2 ? w[2] : w[3]));
  }
};

// Uniform in (0, 1]: ((x >> 8) + 1) * 2^-24, exact in FP32.
__host__ __device__ inline float u01(uint32_t x) {
  return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
}

// IEEE binary32 -> bfloat16, round to nearest even (inputs here are finite).
__host__ __device__ inline uint16_t f32_to_bf16(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint32_t h) {
  return __builtin_bit_cast(float, (uint32_t)(h << 16));
}

// FNV-1a over (player, hole bucket, board bucket, street, history bytes),
// masked to the table (cfr_gpu.cu:165-179).
__host__ __device__ inline uint32_t info_hash(uint32_t player, uint32_t hb, uint32_t bbk,
                                              uint32_t street, uint32_t ab, uint32_t mask) {
  uint32_t h = 2166136261u;
  h = (h ^ (player & 0xFFu)) * 16777619u;
  h = (h ^ (hb & 0xFFu)) * 16777619u;
  h = (h ^ (bbk & 0xFFu)) * 16777619u;
  h = (h ^ (street & 0xFFu)) * 16777619u;
  h = (h ^ (ab & 0xFFu)) * 16777619u;
  h = (h ^ ((ab >> 8) & 0xFFu)) * 16777619u;
  h = (h ^ ((ab >> 16) & 0xFFu)) * 16777619u;
  h = (h ^ ((ab >> 24) & 0xFFu)) * 16777619u;
  return h & mask;
}

// Rank-major 1..52 index of a suit-major card (cfr_gpu.cu:87-90).
__host__ __device__ inline int card_rm(uint32_t c) { return (int)(c % 13u) * 4 + (int)(c / 13u) + 1; }

// Packed (category << 12 | ordinal) -> 1..7462 (cfr_gpu.cu:95-112).
__host__ __device__ inline uint32_t decode_packed(uint32_t packed) {
  const uint32_t cat = (packed >> 12) & 0xFu;
  const uint32_t ord = packed & 0x0FFFu;
  if (ord == 0 || cat < 1 || cat > 9) return 0;
  const uint32_t off = cat == 1 ? 0u : cat == 2 ? 1277u : cat == 3 ? 4137u : cat == 4 ? 4995u
                     : cat == 5 ? 5853u : cat == 6 ? 5863u : cat == 7 ? 7140u : cat == 8 ? 7296u : 7452u;
  return (off + ord) & 0xFFFFu;
}

// Board bucket of a decoded rank, with the device's uint8 truncation (cfr_gpu.cu:488).
__host__ __device__ inline uint32_t bucket_of_rank(uint32_t rank16) {
  return (((uint32_t)((int)rank16 - 1)) * 128u / 7462u) & 0xFFu;
}

// The strategy harness's bucket (abstraction.cpp:61-70): clamped at 127.
__host__ __device__ inline uint32_t eval_bucket_of_rank(uint32_t rank16) {
  const uint32_t b = ((uint32_t)((int)rank16 - 1)) * 128u / 7462u;
  return b < 127u ? b : 127u;
}

// ---------------------------------------------------------------------------
// Betting rules over a menu (cfr_gpu.cu:186-244; abstraction.cpp:144-234).
// ---------------------------------------------------------------------------
__host__ __device__ inline int min_raise_to(int current_bet, int lfr, int bb) {
  return current_bet == 0 ? (bb > 1 ? bb : 1) : current_bet + (lfr > bb ? lfr : bb);
}

__host__ __device__ inline int frac_target(const DevGame& g, int a, int pot, int current_bet, int mrt) {
  int inc = ((int)g.num[a] * pot) / (int)g.den[a];
  inc = inc > 1 ? inc : 1;
  const int t = current_bet + inc;
  return t > mrt ? t : mrt;
}

__host__ __device__ inline uint32_t valid_mask(const DevGame& g, int pot, int stack, int to_call,
                                               int current_bet, int lfr, int raises) {
  if (pot <= 0) pot = 1;
  uint32_t mask;
  if (to_call == 0) {
    mask = g.check_mask;
  } else if (to_call >= stack) {
    return g.fold_mask | g.allin_mask;
  } else {
    mask = g.fold_mask | g.call_mask;
  }
  const int head = stack - to_call;
  if (head <= 0) return mask;
  if (g.max_raises > 0 && raises >= g.max_raises) return mask;
  const int mrt = min_raise_to(current_bet, lfr, g.bb);
  const int maxrt = current_bet + head;
  if (maxrt >= mrt) {
    // Given mrt <= maxrt, a raise's target max(current_bet + max(1, num*pot/den), mrt)
    // fits iff max(1, floor(num*pot/den)) <= head, i.e. num*pot < (head + 1)*den
    // (head >= 1 here): no division needed.
    uint32_t rm = g.raise_mask;
    while (rm) {
      const int a = __builtin_ctz(rm);
      rm &= rm - 1;
      const bool fits = g.kind[a] == MCCFR_ACT_RAISE_MIN || (int)g.num[a] * pot < (head + 1) * (int)g.den[a];
      if (fits) mask |= 1u << a;
    }
  }
  return mask | g.allin_mask;
}

__host__ __device__ inline int chips_for(const DevGame& g, int a, int pot, int stack, int to_call,
                                         int current_bet, int lfr) {
  if (pot <= 0) pot = 1;
  const int kind = g.kind[a];
  if (kind == MCCFR_ACT_FOLD || kind == MCCFR_ACT_CHECK) return 0;
  if (kind == MCCFR_ACT_CALL || kind == MCCFR_ACT_CHECK_CALL) return to_call < stack ? to_call : stack;
  if (kind == MCCFR_ACT_ALL_IN) return stack;
  const int mrt = min_raise_to(current_bet, lfr, g.bb);
  const int head = stack - to_call;
  const int maxrt = current_bet + (head > 0 ? head : 0);
  int target = kind == MCCFR_ACT_RAISE_MIN ? mrt : frac_target(g, a, pot, current_bet, mrt);
  target = target < maxrt ? target : maxrt;
  const int add = target - current_bet;
  const int chips = to_call + (add > 0 ? add : 0);
  return chips < stack ? chips : stack;
}

__host__ __device__ inline int passive_fallback(const DevGame& g, uint32_t mask) {
  for (int i = 0; i < g.n_passive; ++i) {
    const uint32_t a = g.passive_order[i];
    if (mask >> a & 1u) return (int)a;
  }
  return 0;
}

// Builds the device descriptor from the public config (host side).
bool make_dev_game(const mccfr_game_cfg& cfg, DevGame& g, const char** why);

}  // namespace mccfr
