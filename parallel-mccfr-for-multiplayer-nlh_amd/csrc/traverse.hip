// The launch pipeline: the reference's simulate kernel (cfr_gpu.cu:878-1011)
// with its regret-matching and clamp sweeps (:318-351, :1018-1022), for one or
// more tenants at once.  The launch's lane space is a list of segments (one
// tenant each, aligned to kSegAlign lanes, so every wave serves one tenant).
//
//   k_deal        one lane per traversal: Philox deal (around the tenant's
//                 fixed board), hole and board buckets, showdown ranks,
//                 dealer -> compact root records.
//   k_walk        each traversal's root-to-leaf path: both node types follow
//                 exactly one child (cfr_gpu.cu:800, :844), so a traversal is
//                 one path.  Lanes keep the game state in LDS and record the
//                 update-player nodes (key, sampled action, legal mask) and
//                 the terminal utility, and counts each emission block's
//                 records.
//   k_emit_*      records in lane order -> (key, payload) arrays, with the
//                 side-pot payoffs of the traversals that reached a showdown;
//                 the payload carries the record's action word and utility,
//                 so the table update reads records in sorted order without
//                 gathers.
//   k_rs_*        stable LSD radix sort of the records by key
//                 (key = segment << slot_bits | slot; records arrive in
//                 segment order, so the passes need only the slot bits), the
//                 sorted bits spread evenly over ceil(sort_bits / 10) passes.
//   k_runs        every run of equal keys is one table slot of one tenant;
//                 runs are listed (reduce_cand, in k_reduce_rows' leading
//                 workgroups, finds the ends of the long ones).
//                 k_reduce_rows: one lane owns a short run and sums
//                 every action's contributions (cfr_gpu.cu:784-837) in FP64
//                 in record order, then applies them: regrets += sum (CFR+
//                 clamp), strategy sums += sum, BF16 snapshot re-matched.
//                 Runs longer than kShortRun are cut into kPiece-record
//                 pieces (k_reduce_piece, one wave each) whose sums
//                 k_reduce_runs adds in piece order.
// No atomics touch the tables, and every sum has a fixed order, so a tenant's
// result does not depend on which other tenants share its launches.
// Untouched slots need no sweep: clamp and regret matching are per-slot and
// idempotent on them.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernel_util.hpp"
#include "kernels.hpp"

namespace mccfr {
namespace {

constexpr int kDealWG = 256;

__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// Turn byte of a seat's bucket word before the walk has loaded it (the
// device's buckets are at most 198: bucket_of_rank of a decoded rank <= 11547).
constexpr uint32_t kTurnUnknown = 0xFFu;

// sd_state bit of a traversal that ended in a showdown (the low bits hold the folded seats).
constexpr uint32_t kShowdownFlag = 0x80u;

// Record flag above the legal mask (bits 8..23 of a record's action word):
// an external-sampling record stands for one update-player node; its second
// word indexes the node's row of regret deltas v(a) - v(I) (DevScratch::es_rd,
// zero for illegal actions), added to the legal actions' regrets, and the node
// adds its strategy-sum share t * sigma to every legal action.
constexpr uint32_t kEsRecord = 1u << 24;
constexpr uint32_t kLegalBits = 0xFFFFu;

// Waiting lanes (with a next traversal) that trigger a refill round of the walk.
constexpr int kRefillLanes = 48;

// Segment holding `lane` (lane0 ascending; the caller passes a wave-uniform lane).
__device__ __forceinline__ uint32_t find_seg(const DevSeg* __restrict__ segs, uint32_t n_segs, uint32_t lane) {
  uint32_t lo = 0, hi = n_segs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (segs[mid].lane0 <= lane) lo = mid;
    else hi = mid;
  }
  return uniform(lo);
}

// ---------------------------------------------------------------------------
// k_deal.  Buckets come from precomputed tables instead of the reference's
// per-player table walks: the rank-walk word after 5 (6) cards depends only on
// the card set, so the flop (turn) bucket of every 5- (6-) card set is stored
// once, indexed by colex rank (kernels.hip k_build_buckets); the river rank and
// bucket come from the arithmetic evaluator.  Per player: three independent
// loads instead of a chain of dependent gathers into the 130 MB table.
// Preflop root (cfr_gpu.cu:931-937, 963): 2N+5 draws over the full deck, then
// the dealer draw.  Postflop root: the fixed board cards leave the deck, the
// 2N hole cards and the 5 - n_board missing board cards are drawn from the
// remaining deck_size cards, and the dealer is the root's.
// ---------------------------------------------------------------------------
// C(n, k) for n < 52, k <= 6 at [k * 52 + n], built at compile time.
struct BinomTable {
  uint32_t v[7 * 52];
};
constexpr BinomTable make_binom() {
  BinomTable t{};
  for (uint32_t k = 0; k < 7; ++k)
    for (uint32_t n = 0; n < 52; ++n) {
      uint64_t c = n < k ? 0 : 1;
      for (uint32_t j = 0; j < k && n >= k; ++j) c = c * (n - j) / (j + 1);  // exact at every step
      t.v[k * 52 + n] = (uint32_t)c;
    }
  return t;
}
static_assert(make_binom().v[6 * 52 + 51] == 18009460u && make_binom().v[5 * 52 + 51] == 2349060u &&
                  make_binom().v[2 * 52 + 1] == 0u && make_binom().v[0] == 1u,
              "C(51, 6), C(51, 5), C(1, 2), C(0, 0)");
__constant__ BinomTable kBinom = make_binom();

__global__ __launch_bounds__(kDealWG) void k_deal(LaunchCfg lc, const DevSeg* __restrict__ segs,
                                                  const DevGame* __restrict__ tenants, DealInputs in,
                                                  DevScratch sc) {
  __shared__ uint32_t s_lut[2704 / 4];
  __shared__ uint16_t s_off[13 * 8 * 5];
  __shared__ uint32_t s_binom[7 * 52];  // C(n, k) at [k * 52 + n]
  __shared__ uint8_t s_deck[kDeckStride * kDealWG];
  for (int i = threadIdx.x; i < 2704 / 4; i += kDealWG) s_lut[i] = reinterpret_cast<const uint32_t*>(in.lut)[i];
  for (int i = threadIdx.x; i < 13 * 8 * 5; i += kDealWG) s_off[i] = in.at.offsets[i];
  for (int i = threadIdx.x; i < 7 * 52; i += kDealWG) s_binom[i] = kBinom.v[i];
  __syncthreads();
  const uint32_t t = blockIdx.x * kDealWG + threadIdx.x;
  if (t >= lc.lanes) return;
  const uint32_t si = find_seg(segs, lc.n_segs, uniform(t & ~63u));
  const DevSeg& S = segs[si];
  const uint32_t i = t - S.lane0;
  if (i >= S.n) return;  // padding lane of the segment's last wave (the walk clears its record count)
  const DevGame& G = tenants[uniform(S.tenant)];
  const int N = G.N;
  const int nb = G.n_board;

  DrawStream rng;
  rng.init(G.seed_lo, G.seed_hi, S.first_tid + i, G.tag);
  uint8_t* dk = s_deck + threadIdx.x * kDeckStride;
  if (nb == 0) {
    static_assert(kDeckStride % 4 == 0, "dword-aligned decks");
    uint32_t* dw = reinterpret_cast<uint32_t*>(dk);
#pragma unroll
    for (uint32_t q = 0; q < 13; ++q) dw[q] = 0x03020100u + 0x04040404u * q;  // cards 4q .. 4q + 3
  } else {
    const uint64_t bm = ((uint64_t)G.board_hi << 32) | G.board_lo;
    int j = 0;
    for (int k = 0; k < 52; ++k)
      if (!(bm >> k & 1ull)) dk[j++] = (uint8_t)k;
  }
  const uint32_t M = G.deck_size;
  const int n_deal = G.n_deal;
  for (int k = 0; k < n_deal; ++k) {  // partial Fisher-Yates (cfr_gpu.cu:460-470)
    const uint32_t r = rng.next();
    const int j = k + (int)(((uint64_t)r * (uint64_t)(M - (uint32_t)k)) >> 32);
    const uint8_t a = dk[k], b = dk[j];
    dk[k] = b;
    dk[j] = a;
  }
  uint32_t bc[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    int idx = 2 * N + k - nb;
    idx = idx < 0 ? 0 : idx;
    bc[k] = k < nb ? (uint32_t)G.board[k] : (uint32_t)dk[idx];
  }
  const uint8_t* lut = reinterpret_cast<const uint8_t*>(s_lut);
  const size_t L = sc.cap_lanes;
  // every seat's flop-bucket load first (a 2.6 MB table: L2 hits), so all of
  // them are in flight while the ranks are computed.  The turn bucket (a 20 MB
  // table: a sector from memory per load) is left to the walk, which loads it
  // only for the seats that act on the turn; its index is stored here.
  uint32_t bf[kMaxPlayers];
#pragma unroll
  for (int pl = 0; pl < kMaxPlayers; ++pl) {
    bf[pl] = 0;
    if (pl < N) {
      const uint32_t h0 = dk[2 * pl], h1 = dk[2 * pl + 1];
      bf[pl] = in.bt.flop5[colex5_tb(s_binom, h0, h1, bc[0], bc[1], bc[2])];
      sc.root_tidx[(size_t)pl * L + t] = colex6_tb(s_binom, h0, h1, bc[0], bc[1], bc[2], bc[3]);
    }
  }
  Rank7Sums board;  // the five board cards, shared by every seat
#pragma unroll
  for (int k = 0; k < 5; ++k) rank7_add(board, bc[k]);
#pragma unroll
  for (int pl = 0; pl < kMaxPlayers; ++pl) {
    if (pl < N) {
      const uint32_t h0 = dk[2 * pl], h1 = dk[2 * pl + 1];
      const uint32_t hb = lut[h0 * 52 + h1];
      Rank7Sums hand = board;
      rank7_add(hand, h0);
      rank7_add(hand, h1);
      const uint32_t r7 = rank7_finish(hand, in.at.nonflush, in.at.flush, s_off);
      sc.root_bkt[(size_t)pl * L + t] = hb | (bf[pl] << 8) | (kTurnUnknown << 16) | (bucket_of_rank(r7) << 24);
      sc.root_rank[(size_t)pl * L + t] = (uint16_t)r7;
    }
  }
  // draw 2N+5 of the preflop root (cfr_gpu.cu:963)
  sc.root_dealer[t] = G.root_street == 0 ? (uint8_t)(rng.next() % (uint32_t)N) : (uint8_t)G.root_dealer;
}

// A menu entry packed for the walk's LDS copy: kind | num << 8 | den << 16.
__device__ __forceinline__ uint32_t menu_word(const DevGame& g, int a) {
  return (uint32_t)g.kind[a] | ((uint32_t)g.num[a] << 8) | ((uint32_t)g.den[a] << 16);
}

// ceil(2^32 / den) for den >= 2 (0 otherwise): num * pot / den is then one
// high multiply plus one correction (chips_lds).
__device__ __forceinline__ uint32_t den_reciprocal(uint32_t den) { return den >= 2 ? 0xFFFFFFFFu / den + 1u : 0u; }

// chips_for over the wave's LDS copy of the menu (divergent action index);
// rcp[a] = den_reciprocal(den of entry a).  Branch-free: the call, all-in and
// raise amounts are all formed and the action's kind selects one.
__device__ __forceinline__ int chips_lds(const DevGame& g, const uint32_t* menu, const uint32_t* rcp, int a, int pot,
                                         int stack, int to_call, int current_bet, int lfr) {
  pot = pot <= 0 ? 1 : pot;
  const uint32_t m = menu[a];
  const int k = (int)(m & 0xFFu);
  const int call_chips = to_call < stack ? to_call : stack;
  const int mrt = min_raise_to(current_bet, lfr, g.bb);
  const int head = stack - to_call;
  const int maxrt = current_bet + (head > 0 ? head : 0);
  // x / den for x = num * pot < 2^26 and 1 <= den < 2^8: q = hi32(x * ceil(2^32 / den))
  // is x / den or one more (x * e / 2^32 < 1 for e = ceil(2^32 / den) - 2^32 / den < 1);
  // entries other than pot fractions have den 1
  const uint32_t x = __umul24((m >> 8) & 0xFFu, (uint32_t)pot), den = m >> 16;
  uint32_t q = den >= 2 ? __umulhi(x, rcp[a]) : x;
  q = q * den > x ? q - 1u : q;
  const int inc = (int)q > 1 ? (int)q : 1;
  int target = (k == MCCFR_ACT_POT_FRAC && current_bet + inc > mrt) ? current_bet + inc : mrt;
  target = target < maxrt ? target : maxrt;
  const int add = target - current_bet;
  const int raise_chips = min(to_call + (add > 0 ? add : 0), stack);
  return (k == MCCFR_ACT_FOLD || k == MCCFR_ACT_CHECK) ? 0
         : (k == MCCFR_ACT_CALL || k == MCCFR_ACT_CHECK_CALL) ? call_chips
         : k == MCCFR_ACT_ALL_IN ? stack
                                 : raise_chips;
}

// ---------------------------------------------------------------------------
// k_walk.  Wave w's lane l walks the traversals at lanes base + j * jstride + l
// for j < tpl: contiguous (jstride = 64, base = w * 64 * tpl: the wave covers
// 64 * tpl lanes of one segment) or, for single-segment launches, persistent
// (one wave per resident wave slot of the device, base = w * 64, jstride = 64
// times the number of waves).
//
// Per-lane state: seat stacks and bets in LDS (field-major: lane-consecutive
// banks), a bet counts only while its 2-bit street tag matches the street, so
// a street change writes nothing; the action queue is a seat mask plus the
// seat to search from (the reference's queue, cfr_gpu.cu:575-585, lists the
// same seats in the same cyclic order).  The loop only runs decision nodes
// and street changes; showdowns and the result stores run once per round for
// the whole wave.
// ---------------------------------------------------------------------------
constexpr uint32_t kStackMask = 0x7FFFu;  // chips behind: bits 0..14; bet: 15..29; street tag: 30..31

// Betting rules of one tenant held in registers (valid_mask restated over the
// wave's LDS menu; same legality as device_common.hpp valid_mask).
struct Rules {
  uint32_t check_mask, call_mask, fold_mask, allin_mask, raise_mask;
  int max_raises, bb;
  // the first four raise entries, wave-uniform: num | den << 8 | (1 << a) << 16
  // (a min-raise as num 0, den 1: always of legal size); further raise
  // entries are read from the LDS menu
  uint32_t fast[4];
  uint32_t raise_rest;
};

// Branch-free: every lane evaluates every case and the result is selected
// (a wave's lanes sit in different cases; early returns made divergent
// regions).  Same outcomes as valid_mask: facing a bet it cannot cover, a seat
// may fold or go all in; with no chips behind, or at the raise cap, it may only
// check / fold / call; otherwise the raises of legal size and all-in are added.
__device__ __forceinline__ uint32_t rules_mask(const Rules& g, const uint32_t* menu, int pot, int stack, int to_call,
                                               int current_bet, int lfr, int raises) {
  pot = pot <= 0 ? 1 : pot;
  const int head = stack - to_call;
  const uint32_t base = to_call == 0 ? g.check_mask : (g.fold_mask | g.call_mask);
  // max(current_bet + max(1, num*pot/den), mrt) <= current_bet + head
  //   <=> num*pot < (head + 1)*den, given mrt <= current_bet + head
  // (pot < 2^18, head < 2^15, num and den < 2^8: 24-bit multiplies suffice;
  // for head <= 0 the bits are discarded below)
  uint32_t rbits = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t f = g.fast[i];
    const bool ok = __umul24(f & 0xFFu, (uint32_t)pot) < __umul24((uint32_t)(head + 1), (f >> 8) & 0xFFu);
    rbits |= ok ? f >> 16 : 0u;  // an unused slot has no bit
  }
  uint32_t rm = g.raise_rest;  // wave-uniform
  while (rm) {
    const int a = __builtin_ctz(rm);
    rm &= rm - 1;
    const uint32_t m = menu[a];
    const bool ok = (m & 0xFFu) == MCCFR_ACT_RAISE_MIN ||
                    __umul24((m >> 8) & 0xFFu, (uint32_t)pot) < __umul24((uint32_t)(head + 1), m >> 16);
    rbits |= ok ? 1u << a : 0u;
  }
  const bool capped = g.max_raises > 0 && raises >= g.max_raises;
  const int mrt = min_raise_to(current_bet, lfr, g.bb);
  const uint32_t open = (current_bet + head >= mrt ? rbits : 0u) | g.allin_mask;
  const uint32_t mask = (head > 0 && !capped) ? (base | open) : base;
  return (to_call != 0 && to_call >= stack) ? (g.fold_mask | g.allin_mask) : mask;
}

template <int ROW, int WG, bool TRACE>
__global__ __launch_bounds__(WG) void k_walk(LaunchCfg lc, const DevSeg* __restrict__ segs,
                                             const DevGame* __restrict__ tenants, DevScratch sc, uint32_t tpl,
                                             uint32_t jstride,
                                             mccfr_trace* __restrict__ trace) {
  __shared__ uint32_t s_sb[kMaxPlayers * WG];    // stack | bet << 15 | street << 30
  __shared__ uint32_t s_bk[kMaxPlayers * WG];    // buckets: hole | flop << 8 | turn << 16 | river << 24
  __shared__ uint32_t s_ps[kRegPending * WG];    // pending update nodes: key
  __shared__ uint32_t s_pa[kRegPending * WG];    //                       action | legal mask << 8
  __shared__ uint32_t s_menu[WG / 64][MCCFR_MAX_ACTIONS];  // kind | num << 8 | den << 16 of the wave's tenant
  __shared__ uint32_t s_rcp[WG / 64][MCCFR_MAX_ACTIONS];   // den_reciprocal of each entry's den

  const int tx = threadIdx.x, lane = tx & 63, wv = tx >> 6;
  const uint32_t base =
      uniform((uint32_t)((blockIdx.x * WG + tx) >> 6) * 64u * (jstride == 64u ? tpl : 1u));
  if (base >= lc.lanes) return;
  const uint32_t si = find_seg(segs, lc.n_segs, base);
  const DevSeg& Sg = segs[si];
  const DevGame& G = tenants[uniform(Sg.tenant)];
  uint32_t* menu = s_menu[wv];
  uint32_t* rcp = s_rcp[wv];
  if (lane < MCCFR_MAX_ACTIONS) {
    menu[lane] = menu_word(G, lane);
    rcp[lane] = G.kind[lane] == MCCFR_ACT_POT_FRAC ? den_reciprocal(G.den[lane]) : 0u;
  }
  __builtin_amdgcn_wave_barrier();

  uint32_t* sb = s_sb + tx;
  uint32_t* bk = s_bk + tx;
  uint32_t* ps = s_ps + tx;
  uint32_t* pa = s_pa + tx;

  // the tenant's constants, once, in registers
  Rules rules;
  rules.check_mask = G.check_mask;
  rules.call_mask = G.call_mask;
  rules.fold_mask = G.fold_mask;
  rules.allin_mask = G.allin_mask;
  rules.raise_mask = G.raise_mask;
  rules.max_raises = G.max_raises;
  rules.bb = G.bb;
  {
    uint32_t rm = rules.raise_mask;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      rules.fast[i] = 0u;
      if (rm) {
        const int a = __builtin_ctz(rm);
        rm &= rm - 1;
        const bool min_raise = G.kind[a] == MCCFR_ACT_RAISE_MIN;
        rules.fast[i] = (min_raise ? 0u : (uint32_t)G.num[a]) | ((min_raise ? 1u : (uint32_t)G.den[a]) << 8) |
                        ((1u << a) << 16);
      }
    }
    rules.raise_rest = rm;
  }
  const int N = G.N;
  const int hist_bits = G.hist_bits;
  const uint32_t slot_mask = (uint32_t)((1ull << G.table_log2) - 1ull);
  const uint16_t* __restrict__ sig = G.sig;
  const uint32_t seg_key = uniform((uint32_t)((uint64_t)si << lc.slot_bits));
  const uint32_t lane0 = Sg.lane0, seg_n = Sg.n;
  const int up = Sg.update_player;
  const uint32_t allmask = (1u << N) - 1u;
  const size_t L = sc.cap_lanes;

  // The wave's traversals form one pool, t(k) = base + (k / 64) * jstride +
  // k % 64 for k < 64 * tpl, taken in order of k by whichever lanes are free.
  // A lane whose traversal has ended waits until kRefillLanes lanes (or the
  // pool's remaining traversals) are waiting, or until no lane walks; then
  // every waiting lane finishes its traversal (showdown, stores) and takes
  // the pool's next one.  The finishing code runs for many lanes at once, and
  // a lane is idle only once the pool is empty.  A traversal's result depends
  // on t alone (its draws, deal and output slots), not on the lane.
  DrawStream rng;
  uint32_t folded = 0, all_in = 0, ab = 0, qmask = 0, phash = 2166136261u;
  int qpos = 0, pot = 0, current_bet = 0, lfr = rules.bb, street = 0, raises = 0, depth = 0, dealer = 0;
  int npend = 0, decisions = 0, upd_legal = 0;
  int status = 5;  // 0 walking, 1 fold, 2 showdown, 3 depth limit, 5 waiting to start, 6 no traversal left
  float u = 0.0f;
  uint32_t t = 0;
  // the pool's traversals in the segment (t(k) grows with k, so they are the
  // first pool_n); the slots that are a segment's padding lanes hold no records
  uint32_t pool_n = 0;
  for (uint32_t j = 0; j < tpl; ++j) {
    const uint32_t tt = base + j * jstride + (uint32_t)lane;
    const bool in_seg = tt - lane0 < seg_n;
    pool_n += (uint32_t)__popcll(__ballot(in_seg));
    if (tt < lc.lanes && !in_seg) {
      sc.node_count[tt] = 0;
      sc.sd_state[tt] = 0;
    }
  }
  pool_n = uniform(pool_n);
  uint32_t pool_next = 0;  // wave-uniform
  const uint64_t lanes_below = (1ull << lane) - 1ull;
  uint32_t rec_sum = 0;  // records of the lane's finished traversals

  while (true) {
    const uint64_t walking = __ballot(status == 0);
    const uint64_t waiting = __ballot(status != 0 && status != 6);
    if (waiting == 0 && walking == 0) break;
    const uint32_t left = pool_n - pool_next;
    const uint32_t n_wait = (uint32_t)__popcll(waiting);
    if (waiting != 0 &&
        (walking == 0 || (left != 0 && min(n_wait, left) >= min((uint32_t)kRefillLanes, left)))) {
      // The tenant's root and stack fields are read here, where they are used,
      // rather than kept in scalar registers through the walk (the compiler
      // barrier stops their loads from being hoisted out of the loop).
      asm volatile("" ::: "memory");
      // a showdown's payoff is settled by k_emit after the walk: keep
      // every seat's investment and the folded seats
      if (status == 2) {
#pragma unroll
        for (int p = 0; p < kMaxPlayers; ++p)
          if (p < N) sc.sd_inv[(size_t)p * L + t] = (uint16_t)(G.stack[p] - (int)(sb[p * WG] & kStackMask));
      }
      if (status >= 1 && status <= 3) {
        if (status == 3) u = 0.0f;  // depth guard (cfr_gpu.cu:661)
        if (TRACE) {
          mccfr_trace r;
          r.utility = u;
          r.path_hash = phash;
          r.decisions = (uint16_t)decisions;
          r.updates = (uint8_t)npend;
          r.terminal = (uint8_t)status;
          r.update_legal = (uint16_t)upd_legal;
          r.reserved = 0;
          trace[t] = r;
        }
        sc.node_count[t] = (uint8_t)npend;
        // the traversal's records, counted for its emission block: summed per
        // lane and added once per wave after the walk (a wave's traversals lie
        // in one 256-lane block for 1, 2 or 4 per lane), or added at once in
        // the persistent shape
        if (jstride == 64u) rec_sum += (uint32_t)npend;
        else if (npend != 0) atomicAdd(&sc.blk_cnt[t / kEmitLanes], (uint32_t)npend);
        sc.sd_state[t] = status == 2 ? (uint8_t)(kShowdownFlag | folded) : (uint8_t)0;
        if (status != 2) sc.node_u[t] = u;
        const int kr = npend < kRegPending ? npend : kRegPending;
        for (int k = 0; k < kr; ++k) {
          sc.node_slot[(size_t)k * L + t] = ps[k * WG];
          sc.node_av[(size_t)k * L + t] = pa[k * WG];
        }
      }
      const uint32_t k = pool_next + (uint32_t)__popcll(waiting & lanes_below);
      pool_next = uniform(min(pool_n, pool_next + n_wait));
      if (status != 0 && status != 6) {
        status = 6;
        if (k < pool_n) {
          t = base + (k >> 6) * jstride + (k & 63u);
          folded = 0;
          all_in = 0;
          ab = 0;
          qmask = 0;
          phash = 2166136261u;
          qpos = 0;
          pot = 0;
          current_bet = 0;
          lfr = rules.bb;
          street = 0;
          raises = 0;
          depth = 0;
          npend = 0;
          decisions = 0;
          upd_legal = 0;
          u = 0.0f;
          const int root_street = G.root_street;
          // Root state (cfr_gpu.cu:954-981: stacks, blinds, preflop queue from the
          // seat after the big blind; or the tenant's postflop root with the queue
          // from the seat after the dealer, as at a street change, cfr_gpu.cu:676-686).
#pragma unroll
          for (int p = 0; p < kMaxPlayers; ++p) {
            if (p < N) {
              sb[p * WG] = (uint32_t)G.root_stack[p] | ((uint32_t)root_street << 30);
              bk[p * WG] = sc.root_bkt[(size_t)p * L + t];
            }
          }
          dealer = sc.root_dealer[t];
          if (root_street == 0) {
            const int sb_pos = (N == 2) ? dealer : wrap(dealer + 1, N);
            const int bb_pos = (N == 2) ? wrap(dealer + 1, N) : wrap(dealer + 2, N);
            const int s_sb = (int)(sb[sb_pos * WG] & kStackMask);
            const int sb_amt = G.sb < s_sb ? G.sb : s_sb;
            sb[sb_pos * WG] = (uint32_t)(s_sb - sb_amt) | ((uint32_t)sb_amt << 15);
            if (s_sb - sb_amt == 0) all_in |= 1u << sb_pos;
            const int s_bb = (int)(sb[bb_pos * WG] & kStackMask);
            const int bb_amt = rules.bb < s_bb ? rules.bb : s_bb;
            sb[bb_pos * WG] = (uint32_t)(s_bb - bb_amt) | ((uint32_t)bb_amt << 15);
            if (s_bb - bb_amt == 0) all_in |= 1u << bb_pos;
            pot = sb_amt + bb_amt;
            current_bet = bb_amt;
            qmask = allmask & ~all_in;
            qpos = wrap(bb_pos + 1, N);
          } else {
            folded = G.root_folded;
            all_in = G.root_allin;
            pot = G.root_pot;
            street = root_street;
            ab = G.root_ab;
            qmask = allmask & ~(folded | all_in);
            qpos = wrap(dealer + 1, N);
          }
          rng.init(G.seed_lo, G.seed_hi, Sg.first_tid + (t - lane0), G.tag);
          rng.k = G.walk_k0;  // the deal (and the preflop dealer draw) came first
          status = 0;
        }
      }
      continue;
    }
    if (status != 0) continue;
    // one recursion level of traverse_gpu (cfr_gpu.cu:644-860)
    if (depth >= kMaxDepth) {
      status = 3;
      continue;
    }
    const uint32_t alive = allmask & ~folded;
    if (__popc(alive) == 1) {
      // the tenant's fields used on rare paths are read where they are used,
      // not held in scalar registers through the walk
      const int inv_up = G.stack[up] - (int)(sb[up * WG] & kStackMask);
      u = ((int)__builtin_ctz(alive) == up) ? (float)(pot - inv_up) : (float)(-inv_up);
      status = 1;
      continue;
    }
    if (qmask == 0) {
      // street change (cfr_gpu.cu:668-692), then the next level's decision in
      // the same iteration; with nobody left to act, straight to the showdown
      if (street == 3) {
        status = 2;
        continue;
      }
      ++street;
      current_bet = 0;
      lfr = rules.bb;
      raises = 0;
      qmask = allmask & ~(folded | all_in);
      qpos = wrap(dealer + 1, N);
      if (qmask == 0) {
        status = 2;
        continue;
      }
      ++depth;
      if (depth >= kMaxDepth) {
        status = 3;
        continue;
      }
    }

    // decision node: the first queued seat at or after qpos
    const uint32_t rot = ((qmask >> qpos) | (qmask << (N - qpos))) & allmask;
    const int player = wrap(qpos + (int)__builtin_ctz(rot), N);
    const uint32_t sbv = sb[player * WG];
    const int stk = (int)(sbv & kStackMask);
    const int bet = ((int)(sbv >> 30) == street) ? (int)((sbv >> 15) & kStackMask) : 0;
    const int to_call = current_bet - bet;
    const uint32_t vm = rules_mask(rules, menu, pot, stk, to_call, current_bet, lfr, raises);
    uint32_t Bp = bk[player * WG];
    if (street == 2 && ((Bp >> 16) & 0xFFu) == kTurnUnknown) {  // first turn decision of this seat
      const uint32_t tb = sc.turn6[sc.root_tidx[(size_t)player * L + t]];
      Bp = (Bp & 0xFF00FFFFu) | (tb << 16);
      bk[player * WG] = Bp;
    }
    const uint32_t bbk = street == 0 ? 0u : (Bp >> (8 * street)) & 0xFFu;
    const uint32_t slot = info_hash((uint32_t)player, Bp & 0xFFu, bbk, (uint32_t)street, ab, slot_mask);
    float sg[ROW];
    load_bf16_row<ROW>(sig + (size_t)slot * ROW, sg);
    // sample_action (cfr_gpu.cu:510-537) on the BF16 snapshot: running sums
    // of the legal actions' sigma in action order (an illegal action adds 0,
    // which leaves a sum unchanged; legal actions are menu indices below A)
    float cum[ROW];
    float total = 0.0f;
#pragma unroll
    for (int a = 0; a < ROW; ++a) {
      total += (vm >> a & 1u) ? sg[a] : 0.0f;
      cum[a] = total;
    }
    int act = -1;
    if (total >= 1e-7f) {
      // the first legal action whose running sum reaches r.  r > 0 and the
      // sums do not decrease, so that action's index is the number of sums
      // below r (an action that adds nothing is never the first to reach r),
      // and r <= total keeps it below ROW.
      const float r = u01(rng.next()) * total;
      int below = 0;
#pragma unroll
      for (int a = 0; a < ROW; ++a) below += cum[a] < r ? 1 : 0;
      act = below;
    }
    if (act < 0) {  // passive fallback (cfr_gpu.cu:498-508, 520-521)
      for (int i = 0; i < G.n_passive; ++i) {
        const int a = (int)(G.passive_order[i] & 15u);
        if (vm >> a & 1u) {
          act = a;
          break;
        }
      }
      if (act < 0) act = 0;
    }
    if (TRACE) {
      phash = fnv_path(phash, slot, (uint32_t)act);
      ++decisions;
    }
    if (player == up) {
      const uint32_t av = (uint32_t)act | (vm << 8);
      const uint32_t key = seg_key | slot;
      if (npend < kRegPending) {
        ps[npend * WG] = key;
        pa[npend * WG] = av;
      } else {
        sc.node_slot[(size_t)npend * L + t] = key;
        sc.node_av[(size_t)npend * L + t] = av;
      }
      ++npend;
      if (TRACE) upd_legal += __popc(vm);
    }
    // apply_action_gpu + rebuild_queue (cfr_gpu.cu:543-585)
    qmask &= ~(1u << player);
    qpos = wrap(player + 1, N);
    const int chips = chips_lds(G, menu, rcp, act, pot, stk, to_call, current_bet, lfr);
    ab = ((ab << hist_bits) | (uint32_t)act) & kHistoryMask;
    if (rules.fold_mask >> act & 1u) {
      folded |= 1u << player;
    } else {
      const int ns = stk - chips, nb = bet + chips;
      sb[player * WG] = (uint32_t)ns | ((uint32_t)nb << 15) | ((uint32_t)street << 30);
      pot += chips;
      if (ns == 0) all_in |= 1u << player;
      if (nb > current_bet) {
        const int raise = nb - current_bet;
        if (raise >= lfr || current_bet == 0) lfr = raise;
        current_bet = nb;
        ++raises;
        qmask = allmask & ~(folded | all_in) & ~(1u << player);
      }
    }
    ++depth;
  }
  if (jstride == 64u) {  // the wave's records, for its emission block
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) rec_sum += (uint32_t)__shfl_xor((int)rec_sum, d, 64);
    if (lane == 0 && rec_sum != 0) atomicAdd(&sc.blk_cnt[base / kEmitLanes], rec_sum);
  }
}

// Side-pot payoff of the update player at a showdown (cfr_gpu.cu:694-752):
// layers over the distinct investments in ascending order, each shared by the
// best live ranks that reach it.  inv / rkv: every seat's investment and
// 7-card rank (0 beyond the table's seats).
__device__ __forceinline__ float showdown_payoff(int up, uint32_t folded, const int* inv, const uint32_t* rkv) {
  int inv_up = inv[0];
  uint32_t rk_up = rkv[0];
#pragma unroll
  for (int p = 1; p < kMaxPlayers; ++p)
    if (p == up) {
      inv_up = inv[p];
      rk_up = rkv[p];
    }
  const bool up_in = !(folded >> up & 1u);
  float pay = -(float)inv_up;
  // The update player wins exactly the layers above M, the largest investment
  // of a live seat with a better rank (0 if none), up to its own investment;
  // with no live seat of equal rank investing more than M, it wins each of
  // them alone, and their pots add up to sum_p min(inv_p, inv_up) -
  // min(inv_p, M).  Every share is then a whole number of chips below 2^24,
  // so the loop's FP32 sums are exact and equal this one.  A tie takes the
  // loop (split shares round).
  int M = 0;
  bool tie = false;
#pragma unroll
  for (int p = 0; p < kMaxPlayers; ++p) {
    const bool live = !(folded >> p & 1u) && p != up;
    M = (live && rkv[p] > rk_up && inv[p] > M) ? inv[p] : M;
  }
  int won = 0;
#pragma unroll
  for (int p = 0; p < kMaxPlayers; ++p) {
    const bool live = !(folded >> p & 1u) && p != up;
    tie = tie || (live && rkv[p] == rk_up && inv[p] > M);
    won += min(inv[p], inv_up) - min(inv[p], M);
  }
  if (up_in && !tie && M < inv_up) pay += (float)won;
  int prev = 0;
#pragma unroll
  for (int it = 0; it < kMaxPlayers && up_in && tie; ++it) {
    int lvl = 0x7FFFFFFF;
#pragma unroll
    for (int p = 0; p < kMaxPlayers; ++p)
      if (inv[p] > prev && inv[p] < lvl) lvl = inv[p];
    if (lvl != 0x7FFFFFFF) {
      int layer_pot = 0;
      uint32_t best = 0;
#pragma unroll
      for (int p = 0; p < kMaxPlayers; ++p) {
        const int c = (inv[p] < lvl ? inv[p] : lvl) - prev;
        layer_pot += c > 0 ? c : 0;
        if (!(folded >> p & 1u) && inv[p] >= lvl && rkv[p] > best) best = rkv[p];
      }
      if (best > 0) {
        int n_win = 0;
#pragma unroll
        for (int p = 0; p < kMaxPlayers; ++p)
          if (!(folded >> p & 1u) && inv[p] >= lvl && rkv[p] == best) ++n_win;
        if (up_in && inv_up >= lvl && rk_up == best) pay += (float)layer_pot / (float)n_win;
      }
      prev = lvl;
    }
  }
  return pay;
}

// ---------------------------------------------------------------------------
// k_walk_es: external sampling (mccfr.h MCCFR_SAMPLE_EXTERNAL).  One lane per
// traversal walks its tree depth first: other seats' nodes sample one action
// from the BF16 snapshot as in k_walk, and an update-player node explores
// every legal action in action order.  Such a node pushes a frame (the state
// before its action, its slot and legal mask, the actions left and the value
// of each explored child) onto the lane's stack in HBM, laid out
// [lane][frame][word] (16-byte accesses; a wave's lanes sit at different
// depths, so lane-interleaved words would not coalesce either); a leaf's
// value returns into the top frame, which restores its state for the next
// action or, with none left, forms v(I) = sum_a sigma(a) v(a) (sigma from the
// FP32 regrets, as the reference's update node forms it) and emits one
// record (mask << 8 | kEsRecord, its index) with its row of v(a) - v(I) in
// DevScratch::es_rd at that index.  The update player acts at most on every
// other level, so the depth guard of 20 bounds the stack at kEsFrames frames.
// Every update-player action's subtree is a task with a stream of its own
// (DrawStream::enter_child: the traversal's counters with the task's path
// code, from draw 0), so a subtree's draws do not depend on the order in which
// subtrees are walked; the trace's path hash sums each entered (node, action)
// pair's FNV-1a and its terminal field holds the set of leaf kinds reached.
// Records are appended to the sort input in per-wave reservations of
// kEsChunk (one atomic each; k_es_begin zeroes the counter, unused tails hold
// null records); beyond the input's capacity nothing more is written and k_emit_scan marks
// the launch as overflowed (no table is touched).  Showdown payoffs are
// settled where the walk reaches them, since the value is needed at once.
// ---------------------------------------------------------------------------
constexpr int kEsFrames = 10;
// Traversals per lane of the external-sampling walk (a wave's lanes share its
// 64 * kEsTpl traversals; 64 * kEsTpl <= 256 keeps them in one segment).
constexpr uint32_t kEsTpl = 4;
static_assert(64 * kEsTpl <= kSegAlign && kSegAlign % (64 * kEsTpl) == 0, "a wave's traversals lie in one segment");
// Workgroups of the external-sampling walk for a launch of `lanes` lanes.
inline uint32_t es_grid(uint32_t lanes) { return (lanes + 256u * kEsTpl - 1u) / (256u * kEsTpl); }
template <int ROW>
constexpr int es_frame_words() {
  return 16 + ROW;  // 0..5 seats, 6 seat masks + queue, 7 history, 8 pot, 9 bet, 10 lfr, 11 street | raises |
                    // depth, 12 slot, 13 legal | left << 16, 14 the node's task path, 15 unused, 16.. child values
}

__global__ void k_es_begin(DevScratch sc) {
  if (threadIdx.x == 0) {
    sc.cnt->n_es = 0;
    sc.cnt->n_es_valid = 0;
  }
}

// Records of one reservation of k_walk_es: at least the most one iteration of
// a wave can emit (a node per lane).
constexpr uint32_t kEsChunk = 256;

template <int ROW, bool TRACE>
__global__ __launch_bounds__(256) void k_walk_es(LaunchCfg lc, const DevSeg* __restrict__ segs,
                                                 const DevGame* __restrict__ tenants, DevScratch sc,
                                                 uint32_t* __restrict__ stack, mccfr_trace* __restrict__ trace) {
  constexpr int WG = 256;
  constexpr int FW = es_frame_words<ROW>();
  __shared__ uint32_t s_sb[kMaxPlayers * WG];
  __shared__ uint32_t s_bk[kMaxPlayers * WG];
  __shared__ uint32_t s_menu[WG / 64][MCCFR_MAX_ACTIONS];
  __shared__ uint32_t s_rcp[WG / 64][MCCFR_MAX_ACTIONS];
  const int tx = threadIdx.x, lane = tx & 63, wv = tx >> 6;
  // the wave's kEsTpl * 64 traversals, one 256-lane block of one segment; its
  // lanes take them one after another, a lane taking the next as soon as its
  // tree is done (P: the lane's slot among all the launch's lanes, for its stack)
  const uint32_t P = blockIdx.x * WG + (uint32_t)tx;
  const uint32_t base = uniform((P >> 6) * 64u * kEsTpl);
  if (base >= lc.lanes) return;
  const uint32_t range_end = min(base + 64u * kEsTpl, lc.lanes);
  const uint32_t si = find_seg(segs, lc.n_segs, base);
  const DevSeg& Sg = segs[si];
  const DevGame& G = tenants[uniform(Sg.tenant)];
  uint32_t* menu = s_menu[wv];
  uint32_t* rcp = s_rcp[wv];
  if (lane < MCCFR_MAX_ACTIONS) {
    menu[lane] = menu_word(G, lane);
    rcp[lane] = G.kind[lane] == MCCFR_ACT_POT_FRAC ? den_reciprocal(G.den[lane]) : 0u;
  }
  __builtin_amdgcn_wave_barrier();
  uint32_t* sb = s_sb + tx;
  uint32_t* bk = s_bk + tx;
  Rules rules;
  rules.check_mask = G.check_mask;
  rules.call_mask = G.call_mask;
  rules.fold_mask = G.fold_mask;
  rules.allin_mask = G.allin_mask;
  rules.raise_mask = G.raise_mask;
  rules.max_raises = G.max_raises;
  rules.bb = G.bb;
  {
    uint32_t rm = rules.raise_mask;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      rules.fast[i] = 0u;
      if (rm) {
        const int a = __builtin_ctz(rm);
        rm &= rm - 1;
        const bool min_raise = G.kind[a] == MCCFR_ACT_RAISE_MIN;
        rules.fast[i] = (min_raise ? 0u : (uint32_t)G.num[a]) | ((min_raise ? 1u : (uint32_t)G.den[a]) << 8) |
                        ((1u << a) << 16);
      }
    }
    rules.raise_rest = rm;
  }
  const int N = G.N;
  const int A = G.num_actions;
  const int hist_bits = G.hist_bits;
  const uint32_t slot_mask = (uint32_t)((1ull << G.table_log2) - 1ull);
  const uint32_t seg_key = uniform((uint32_t)((uint64_t)si << lc.slot_bits));
  const int up = Sg.update_player;
  const uint32_t allmask = (1u << N) - 1u;
  const size_t L = sc.cap_lanes;   // root records
  const uint32_t cap_rec = sc.cap_lanes * (uint32_t)kMaxDepth;

  uint32_t t = 0;          // the lane's traversal (its lane in the launch's lane space)
  uint32_t folded = 0, all_in = 0, ab = 0, qmask = 0, phash = 2166136261u;
  int qpos = 0, pot = 0, current_bet = 0, lfr = rules.bb, street = 0, raises = 0, depth = 0, dealer = 0;
  DrawStream rng;
  // root state of traversal t (as k_walk)
  auto start = [&]() {
    folded = 0;
    all_in = 0;
    ab = 0;
    qmask = 0;
    phash = 2166136261u;
    qpos = 0;
    pot = 0;
    current_bet = 0;
    lfr = rules.bb;
    street = 0;
    raises = 0;
    depth = 0;
    dealer = (int)sc.root_dealer[t];
    const int root_street = G.root_street;
#pragma unroll
    for (int p = 0; p < kMaxPlayers; ++p) {
      if (p < N) {
        sb[p * WG] = (uint32_t)G.root_stack[p] | ((uint32_t)root_street << 30);
        bk[p * WG] = sc.root_bkt[(size_t)p * L + t];
      }
    }
    if (root_street == 0) {
      const int sb_pos = (N == 2) ? dealer : wrap(dealer + 1, N);
      const int bb_pos = (N == 2) ? wrap(dealer + 1, N) : wrap(dealer + 2, N);
      const int s_sb = (int)(sb[sb_pos * WG] & kStackMask);
      const int sb_amt = G.sb < s_sb ? G.sb : s_sb;
      sb[sb_pos * WG] = (uint32_t)(s_sb - sb_amt) | ((uint32_t)sb_amt << 15);
      if (s_sb - sb_amt == 0) all_in |= 1u << sb_pos;
      const int s_bb = (int)(sb[bb_pos * WG] & kStackMask);
      const int bb_amt = rules.bb < s_bb ? rules.bb : s_bb;
      sb[bb_pos * WG] = (uint32_t)(s_bb - bb_amt) | ((uint32_t)bb_amt << 15);
      if (s_bb - bb_amt == 0) all_in |= 1u << bb_pos;
      pot = sb_amt + bb_amt;
      current_bet = bb_amt;
      qmask = allmask & ~all_in;
      qpos = wrap(bb_pos + 1, N);
    } else {
      folded = G.root_folded;
      all_in = G.root_allin;
      pot = G.root_pot;
      street = root_street;
      ab = G.root_ab;
      qmask = allmask & ~(folded | all_in);
      qpos = wrap(dealer + 1, N);
    }
    rng.init(G.seed_lo, G.seed_hi, Sg.first_tid + (t - Sg.lane0), G.tag);
    rng.k = G.walk_k0;
  };

  int sp = 0;            // frames on the stack
  bool done = true;      // no traversal in hand (its root value is known, or none taken yet)
  uint32_t next = base;  // the wave's next traversal to hand out (wave uniform)
  bool ret_mode = false; // a value is returning up the tree
  float ret = 0.0f;
  int decisions = 0, updates = 0, upd_legal = 0, leaves = 0, terminal = 0;
  // a lane's frames lie together, [lane][frame][word] (the lanes of a wave sit
  // at different depths, so per-lane runs of 16-byte accesses beat
  // lane-interleaved words)
  auto fr = [&](int f, int w) -> uint32_t& { return stack[((size_t)P * kEsFrames + f) * FW + w]; };
  auto fr4 = [&](int f, int q) -> uint4& { return reinterpret_cast<uint4*>(&fr(f, 0))[q]; };

  // apply_action_gpu + rebuild_queue (cfr_gpu.cu:543-585) for `player`
  auto apply = [&](int player, int act) {
    const uint32_t sbv = sb[player * WG];
    const int stk = (int)(sbv & kStackMask);
    const int bet = ((int)(sbv >> 30) == street) ? (int)((sbv >> 15) & kStackMask) : 0;
    const int to_call = current_bet - bet;
    qmask &= ~(1u << player);
    qpos = wrap(player + 1, N);
    const int chips = chips_lds(G, menu, rcp, act, pot, stk, to_call, current_bet, lfr);
    ab = ((ab << hist_bits) | (uint32_t)act) & kHistoryMask;
    if (rules.fold_mask >> act & 1u) {
      folded |= 1u << player;
    } else {
      const int ns = stk - chips, nb = bet + chips;
      sb[player * WG] = (uint32_t)ns | ((uint32_t)nb << 15) | ((uint32_t)street << 30);
      pot += chips;
      if (ns == 0) all_in |= 1u << player;
      if (nb > current_bet) {
        const int raise = nb - current_bet;
        if (raise >= lfr || current_bet == 0) lfr = raise;
        current_bet = nb;
        ++raises;
        qmask = allmask & ~(folded | all_in) & ~(1u << player);
      }
    }
    ++depth;
  };

  // the wave's current reservation [c_cur, c_end) of the sort input (wave
  // uniform); kEsDead once the input is full
  constexpr uint32_t kEsDead = 0xFFFFFFFFu;
  uint32_t c_cur = 0, c_end = 0, valid = 0;
  auto close_chunk = [&]() {  // null records in the reservation's unused tail
    if (c_cur == kEsDead) return;
    for (uint32_t k = c_cur + (uint32_t)lane; k < c_end; k += 64) {
      sc.key[0][k] = lc.null_key;
      sc.val[0][k] = make_uint2(0u, 0u);
    }
    c_cur = c_end;
  };
  while (true) {
    // lanes without a traversal take the wave's next ones (padding lanes of
    // the segment are skipped: they walk nothing)
    const uint64_t idle = __ballot(done);
    if (idle && next < range_end) {
      const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
      if (done && next + rank < range_end) {
        t = next + rank;
        if (t - Sg.lane0 < Sg.n) {
          start();
          done = false;
          sp = 0;
          ret_mode = false;
          ret = 0.0f;
          decisions = updates = upd_legal = leaves = terminal = 0;
        }
      }
      next = min(next + (uint32_t)__popcll(idle), range_end);
    }
    if (__ballot(!done) == 0 && next >= range_end) break;
    uint32_t emit_k = 0, e_slot = 0, e_vm = 0;
    int ap_player = -1, ap_act = 0;  // the action both paths apply, once, below
    float rd[ROW];
#pragma unroll
    for (int a = 0; a < ROW; ++a) rd[a] = 0.0f;
    if (!done && !ret_mode) {
      // one recursion level of traverse_gpu (cfr_gpu.cu:644-860)
      bool leaf = false;
      if (depth >= kMaxDepth) {
        ret = 0.0f;  // depth guard (cfr_gpu.cu:661)
        terminal |= 1 << 3;
        leaf = true;
      } else if (__popc(allmask & ~folded) == 1) {
        const int inv_up = G.stack[up] - (int)(sb[up * WG] & kStackMask);
        ret = ((int)__builtin_ctz(allmask & ~folded) == up) ? (float)(pot - inv_up) : (float)(-inv_up);
        terminal |= 1 << 1;
        leaf = true;
      } else if (qmask == 0) {
        // street change (cfr_gpu.cu:668-692); with nobody left to act, or
        // after the river, the showdown
        bool showdown = street == 3;
        if (!showdown) {
          ++street;
          current_bet = 0;
          lfr = rules.bb;
          raises = 0;
          qmask = allmask & ~(folded | all_in);
          qpos = wrap(dealer + 1, N);
          showdown = qmask == 0;
          if (!showdown) ++depth;
        }
        if (showdown) {
          int inv[kMaxPlayers];
          uint32_t rkv[kMaxPlayers];
#pragma unroll
          for (int p = 0; p < kMaxPlayers; ++p) {
            inv[p] = p < N ? G.stack[p] - (int)(sb[p * WG] & kStackMask) : 0;
            rkv[p] = p < N ? (uint32_t)sc.root_rank[(size_t)p * L + t] : 0u;
          }
          ret = showdown_payoff(up, folded, inv, rkv);
          terminal |= 1 << 2;
          leaf = true;
        }
      } else {
        const uint32_t rot = ((qmask >> qpos) | (qmask << (N - qpos))) & allmask;
        const int player = wrap(qpos + (int)__builtin_ctz(rot), N);
        const uint32_t sbv = sb[player * WG];
        const int stk = (int)(sbv & kStackMask);
        const int bet = ((int)(sbv >> 30) == street) ? (int)((sbv >> 15) & kStackMask) : 0;
        const int to_call = current_bet - bet;
        const uint32_t vm = rules_mask(rules, menu, pot, stk, to_call, current_bet, lfr, raises);
        uint32_t Bp = bk[player * WG];
        if (street == 2 && ((Bp >> 16) & 0xFFu) == kTurnUnknown) {  // first turn decision of this seat
          const uint32_t tb = sc.turn6[sc.root_tidx[(size_t)player * L + t]];
          Bp = (Bp & 0xFF00FFFFu) | (tb << 16);
          bk[player * WG] = Bp;
        }
        const uint32_t bbk = street == 0 ? 0u : (Bp >> (8 * street)) & 0xFFu;
        const uint32_t slot = info_hash((uint32_t)player, Bp & 0xFFu, bbk, (uint32_t)street, ab, slot_mask);
        ++decisions;
        int act = -1;
        if (player == up && sp >= kEsFrames) {
          // cannot happen (the update player acts on at most every other
          // level); flagged for the host rather than written out of bounds
          sc.cnt->es_overflow = 2u;
          ret = 0.0f;
          terminal |= 1 << 3;
          leaf = true;
        } else if (player == up) {
          // push the node: its state before the action, slot, legal actions
          ++updates;
          upd_legal += __popc(vm);
          uint32_t w[6];
#pragma unroll
          for (int p = 0; p < kMaxPlayers; ++p) w[p] = p < N ? sb[p * WG] : 0u;
          fr4(sp, 0) = make_uint4(w[0], w[1], w[2], w[3]);
          fr4(sp, 1) = make_uint4(w[4], w[5], folded | (all_in << 8) | (qmask << 16) | ((uint32_t)qpos << 24), ab);
          fr4(sp, 2) = make_uint4((uint32_t)pot, (uint32_t)current_bet, (uint32_t)lfr,
                                  (uint32_t)street | ((uint32_t)raises << 8) | ((uint32_t)depth << 16));
          act = __builtin_ctz(vm);
          fr4(sp, 3) = make_uint4(slot, vm | ((vm & (vm - 1u)) << 16), rng.path, 0u);
          ++sp;
          rng.enter_child(rng.path, act);  // the first action's task
        } else {
          // sample_action (cfr_gpu.cu:510-537) on the BF16 snapshot, as k_walk
          float sg[ROW];
          load_bf16_row<ROW>(G.sig + (size_t)slot * ROW, sg);
          float cum[ROW];
          float total = 0.0f;
#pragma unroll
          for (int a = 0; a < ROW; ++a) {
            total += (vm >> a & 1u) ? sg[a] : 0.0f;
            cum[a] = total;
          }
          if (total >= 1e-7f) {
            const float r = u01(rng.next()) * total;
            int below = 0;
#pragma unroll
            for (int a = 0; a < ROW; ++a) below += cum[a] < r ? 1 : 0;
            act = below;
          }
          if (act < 0) {  // passive fallback (cfr_gpu.cu:498-508)
            for (int i = 0; i < G.n_passive; ++i) {
              const int a = (int)(G.passive_order[i] & 15u);
              if (vm >> a & 1u) {
                act = a;
                break;
              }
            }
            if (act < 0) act = 0;
          }
        }
        if (act >= 0) {
          if (TRACE) phash += fnv_path(2166136261u, slot, (uint32_t)act);
          ap_player = player;
          ap_act = act;
        }
      }
      if (leaf) {
        ++leaves;
        ret_mode = true;
      }
    }
    // A leaf's value returns in the iteration that reached it (the forward
    // and return paths run side by side in a wave anyway; a leaf applies no
    // action).  A completed node's value goes on into its parent in the same
    // iteration when that emits nothing (the parent has an action left, or the
    // tree is done): each iteration applies at most one action and emits at
    // most one node's records.
#pragma unroll 1
    for (int pass = 0; pass < 2 && !done && ret_mode; ++pass) {
      // the value `ret` returns into the top frame
      if (sp == 0) {
        done = true;
        if (TRACE) {
          mccfr_trace r;
          r.utility = ret;
          r.path_hash = phash;
          r.decisions = (uint16_t)decisions;
          r.updates = (uint8_t)updates;
          r.terminal = (uint8_t)terminal;
          r.update_legal = (uint16_t)upd_legal;
          r.reserved = (uint16_t)leaves;
          trace[t] = r;
        }
      } else {
        const int f = sp - 1;
        const uint4 q3 = fr4(f, 3);
        const uint32_t w13 = q3.y;
        const uint32_t vm = w13 & 0xFFFFu, left = w13 >> 16;
        if (pass == 1 && !left) break;  // a second completion waits for the next iteration
        const int cur = 31 - (int)__builtin_clz(vm & ~left);  // explored in action order
        fr(f, 16 + cur) = __float_as_uint(ret);
        if (left) {
          // the node's next action, from its saved state
          const int a = __builtin_ctz(left);
          fr(f, 13) = vm | ((left & (left - 1u)) << 16);
          const uint4 q0 = fr4(f, 0), q1 = fr4(f, 1), q2 = fr4(f, 2);
          const uint32_t w[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
#pragma unroll
          for (int p = 0; p < kMaxPlayers; ++p)
            if (p < N) sb[p * WG] = w[p];
          const uint32_t w6 = q1.z;
          folded = w6 & 0xFFu;
          all_in = (w6 >> 8) & 0xFFu;
          qmask = (w6 >> 16) & 0xFFu;
          qpos = (int)(w6 >> 24);
          ab = q1.w;
          pot = (int)q2.x;
          current_bet = (int)q2.y;
          lfr = (int)q2.z;
          const uint32_t w11 = q2.w;
          street = (int)(w11 & 0xFFu);
          raises = (int)((w11 >> 8) & 0xFFu);
          depth = (int)(w11 >> 16);
          if (TRACE) phash += fnv_path(2166136261u, q3.x, (uint32_t)a);
          rng.enter_child(q3.z, a);  // the action's task
          ap_player = up;
          ap_act = a;
          ret_mode = false;
        } else {
          // every action explored: v(I) and the node's records
          const uint32_t slot = q3.x;
          float rr[ROW], v[ROW];
          load_f32_row<ROW>(G.R + (size_t)slot * 2 * ROW, rr);
          float sum_r = 0.0f;
#pragma unroll
          for (int a = 0; a < ROW; ++a) {
            const bool legal = a < A && (vm >> a & 1u);
            rr[a] = legal ? fmaxf(0.0f, rr[a]) : 0.0f;
            sum_r += rr[a];
            v[a] = legal ? __uint_as_float(fr(f, 16 + a)) : 0.0f;
          }
          const float uni = 1.0f / (float)__popc(vm);
          float vi = 0.0f;
#pragma unroll
          for (int a = 0; a < ROW; ++a) {
            const bool legal = a < A && (vm >> a & 1u);
            const float sig = sum_r > 1e-7f ? rr[a] / sum_r : (legal ? uni : 0.0f);
            if (legal) vi += sig * v[a];
          }
#pragma unroll
          for (int a = 0; a < ROW; ++a) rd[a] = v[a] - vi;
          emit_k = 1u;
          e_slot = slot;
          e_vm = vm;
          ret = vi;
          --sp;
        }
      }
    }
    if (ap_player >= 0) apply(ap_player, ap_act);
    // records of the nodes completed in this iteration: a wave prefix over
    // the lanes' counts, written into the wave's reservation (a new one of
    // kEsChunk records, one atomic, when this iteration's do not fit)
    uint32_t incl = emit_k;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(incl, d, 64);
      if (lane >= d) incl += o;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    if (total) {
      if (c_cur != kEsDead && c_end - c_cur < total) {
        close_chunk();
        // once the input is full the counter stops growing (it cannot wrap)
        // and the launch is flagged; nothing more is written
        uint32_t base = 0;
        if (lane == 0) {
          if (__hip_atomic_load(&sc.cnt->n_es, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > cap_rec)
            base = kEsDead;
          else
            base = atomicAdd(&sc.cnt->n_es, kEsChunk);
        }
        base = __shfl(base, 0, 64);
        if (base == kEsDead || base + kEsChunk > cap_rec) {
          if (lane == 0) sc.cnt->es_overflow = 1u;
          c_cur = c_end = kEsDead;
        } else {
          c_cur = base;
          c_end = base + kEsChunk;
        }
      }
      if (c_cur != kEsDead) {
        if (emit_k) {
          // the node's record and its row of regret deltas (16-byte stores)
          const uint32_t off = c_cur + incl - emit_k;
          sc.key[0][off] = seg_key | e_slot;
          sc.val[0][off] = make_uint2((e_vm << 8) | kEsRecord, off);
          f32x4* row = reinterpret_cast<f32x4*>(sc.es_rd + (size_t)off * ROW);
#pragma unroll
          for (int q4 = 0; q4 < ROW / 4; ++q4) {
            f32x4 o;
            o.x = (e_vm >> (4 * q4) & 1u) ? rd[4 * q4] : 0.0f;
            o.y = (e_vm >> (4 * q4 + 1) & 1u) ? rd[4 * q4 + 1] : 0.0f;
            o.z = (e_vm >> (4 * q4 + 2) & 1u) ? rd[4 * q4 + 2] : 0.0f;
            o.w = (e_vm >> (4 * q4 + 3) & 1u) ? rd[4 * q4 + 3] : 0.0f;
            row[q4] = o;
          }
        }
        c_cur += total;
      }
      valid += total;
    }
  }
  close_chunk();
  if (lane == 0 && valid) atomicAdd(&sc.cnt->n_es_valid, valid);
}

// Exclusive prefix of v over a 256-thread workgroup (every thread calls it);
// the workgroup total in *total when total is not null.
__device__ __forceinline__ uint32_t block_excl_scan_256(uint32_t v, uint32_t* s_w, uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) s_w[w] = incl;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t x = s_w[k];
    if (k < w) before += x;
    tot += x;
  }
  __syncthreads();
  if (total) *total = tot;
  return before + incl - v;
}

// ---------------------------------------------------------------------------
// Record emission: lane-ordered (key, (action | mask << 8, utility)) pairs.
// Blocks of kEmitLanes lanes, one lane per thread.
// ---------------------------------------------------------------------------

// Exclusive scan of the block counts in place (one workgroup; the walk added
// the counts); resets the launch counters and fills the per-segment update
// context of the table update.
__global__ __launch_bounds__(1024) void k_emit_scan(LaunchCfg lc, const DevSeg* __restrict__ segs,
                                                    const DevGame* __restrict__ tenants, DevScratch sc) {
  __shared__ uint32_t s_w[16];
  if (lc.es) {
    // external sampling: k_walk_es appended the records itself; a launch whose
    // records exceeded the sort input updates no table and is flagged
    for (uint32_t i = threadIdx.x; i < lc.n_segs; i += blockDim.x) {
      const DevSeg& S = segs[i];
      const DevGame& G = tenants[S.tenant];
      SegCtx c;
      c.R = G.R;
      c.S = G.S;
      c.sig = G.sig;
      c.tw = (float)S.t_weight;
      c.num_actions = (int16_t)G.num_actions;
      c.cfr_plus = (int16_t)(G.use_cfr_plus != 0);
      sc.seg_ctx[i] = c;
    }
    for (uint32_t i = threadIdx.x; i < head1_bytes(sc.cap_lanes) / 4; i += blockDim.x)
      reinterpret_cast<uint32_t*>(sc.head1)[i] = 0x40404040u;
    if (threadIdx.x == 0) {
      // the walk reserved n_es records in whole chunks (the tails of which it
      // filled with null records, keyed above every real key, so they sort
      // last) and emitted n_es_valid real ones
      const bool bad = sc.cnt->es_overflow != 0u;  // sticky until the host has seen it
      sc.cnt->n_sort = bad ? 0u : sc.cnt->n_es;
      sc.cnt->n_rec = bad ? 0u : sc.cnt->n_es_valid;
      sc.cnt->n_long = 0;
      sc.cnt->n_pieces = 0;
      sc.cnt->n_cand = 0;
      sc.cnt->n_runs = 0;
    }
    return;
  }
  for (uint32_t i = threadIdx.x; i < lc.n_segs; i += blockDim.x) {
    const DevSeg& S = segs[i];
    const DevGame& G = tenants[S.tenant];
    SegCtx c;
    c.R = G.R;
    c.S = G.S;
    c.sig = G.sig;
    c.tw = (float)S.t_weight;
    c.num_actions = (int16_t)G.num_actions;
    c.cfr_plus = (int16_t)(G.use_cfr_plus != 0);
    sc.seg_ctx[i] = c;
  }
  const uint32_t nblk = (lc.lanes + kEmitLanes - 1) / kEmitLanes;
  const uint32_t per = (nblk + 1023) / 1024;
  const uint32_t b0 = min(threadIdx.x * per, nblk), b1 = min(b0 + per, nblk);
  // up to kHeld counts per thread are loaded together into registers (launches
  // of up to 2^22 lanes); longer ranges loop
  constexpr uint32_t kHeld = 16;
  const bool held = per <= kHeld;  // workgroup-uniform
  uint32_t hv[kHeld];
  uint32_t sum = 0;
  if (held) {
#pragma unroll
    for (uint32_t k = 0; k < kHeld; ++k) {
      hv[k] = b0 + k < b1 ? sc.blk_cnt[b0 + k] : 0u;
      sum += hv[k];
    }
  } else {
    for (uint32_t b = b0; b < b1; ++b) sum += sc.blk_cnt[b];
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) s_w[w] = incl;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (int k = 0; k < 16; ++k) {
    const uint32_t x = s_w[k];
    if (k < w) before += x;
    tot += x;
  }
  uint32_t run = before + incl - sum;
  if (held) {
#pragma unroll
    for (uint32_t k = 0; k < kHeld; ++k) {
      if (b0 + k < b1) sc.blk_cnt[b0 + k] = run;
      run += hv[k];
    }
  } else {
    for (uint32_t b = b0; b < b1; ++b) {
      const uint32_t c = sc.blk_cnt[b];
      sc.blk_cnt[b] = run;
      run += c;
    }
  }
  // k_runs marks the 64-region groups that hold a run head
  for (uint32_t i = threadIdx.x; i < head1_bytes(sc.cap_lanes) / 4; i += blockDim.x)
    reinterpret_cast<uint32_t*>(sc.head1)[i] = 0x40404040u;
  if (threadIdx.x == 0) {
    sc.cnt->n_rec = tot;
    sc.cnt->n_sort = tot;
    sc.cnt->n_long = 0;
    sc.cnt->n_pieces = 0;
    sc.cnt->n_cand = 0;
    sc.cnt->n_runs = 0;
  }
}

// A block's records are staged in LDS (when they fit) and written out
// contiguously, instead of each lane storing its own records.  One lane per
// thread, so every read of the node lists ([node][lane]) is coalesced.  The
// walk added every block's record count to blk_cnt and k_emit_scan turned the
// counts into offsets; a block takes its offset and clears the entry for the
// next launch's counts.  A traversal that ended in a showdown gets its payoff
// here (the walk left every seat's investment and the folded seats); the
// others carry the walk's utility.  The stage holds 1,024 records (4 per lane;
// a block with more writes its records directly), so 12 KB of LDS leave room
// for 8 waves per SIMD (a 2,048-record stage allowed 6: 23.6 -> 22.2 us).
constexpr uint32_t kEmitStage = 1024;
template <bool TRACE>
__global__ __launch_bounds__(256) void k_emit(LaunchCfg lc, const DevSeg* __restrict__ segs,
                                              const DevGame* __restrict__ tenants, DevScratch sc,
                                              mccfr_trace* __restrict__ trace) {
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_base;
  __shared__ uint32_t s_key[kEmitStage];
  __shared__ uint2 s_val[kEmitStage];
  const uint32_t t = blockIdx.x * kEmitLanes + threadIdx.x;
  const bool in = t < lc.lanes;
  const uint32_t c = in ? (uint32_t)sc.node_count[t] : 0u;
  const uint32_t st = in ? (uint32_t)sc.sd_state[t] : 0u;
  // payoffs of the showdowns that emit records (traced launches: of all)
  const bool showdown = (st & kShowdownFlag) != 0 && (TRACE || c != 0);
  const size_t L = sc.cap_lanes;
  // every load is issued before the block scan (its barriers), so their
  // latencies overlap it: a lane's first kAhead records (most lanes hold at
  // most kAhead), its utility or the showdown's investments and ranks
  constexpr uint32_t kAhead = 4;
  uint32_t ak[kAhead], aa[kAhead];
#pragma unroll
  for (uint32_t k = 0; k < kAhead; ++k) {
    ak[k] = k < c ? sc.node_slot[(size_t)k * L + t] : 0u;
    aa[k] = k < c ? sc.node_av[(size_t)k * L + t] : 0u;
  }
  int N = 0, up = 0;
  if (uniform(t & ~63u) < lc.lanes) {  // wave-uniform: the wave's segment
    const DevSeg& Sg = segs[find_seg(segs, lc.n_segs, uniform(t & ~63u))];
    N = tenants[uniform(Sg.tenant)].N;
    up = Sg.update_player;
  }
  int inv[kMaxPlayers];
  uint32_t rkv[kMaxPlayers];
#pragma unroll
  for (int p = 0; p < kMaxPlayers; ++p) {
    inv[p] = showdown && p < N ? (int)sc.sd_inv[(size_t)p * L + t] : 0;
    rkv[p] = showdown && p < N ? (uint32_t)sc.root_rank[(size_t)p * L + t] : 0u;
  }
  uint32_t ub = c && !showdown ? __float_as_uint(sc.node_u[t]) : 0u;
  if (threadIdx.x == 0) {
    s_base = sc.blk_cnt[blockIdx.x];
    sc.blk_cnt[blockIdx.x] = 0u;  // (this thread's own load of it comes first)
  }
  uint32_t total = 0;
  uint32_t pos = block_excl_scan_256(c, s_w, &total);
  const uint32_t gbase = s_base;
  if (showdown) {
    const float pay = showdown_payoff(up, st & ~kShowdownFlag, inv, rkv);
    ub = __float_as_uint(pay);
    if (TRACE) trace[t].utility = pay;
  }
  const bool staged = total <= kEmitStage;  // block-uniform
  auto put = [&](uint32_t key, uint2 v) {
    if (staged) {
      s_key[pos] = key;
      s_val[pos] = v;
    } else {
      sc.key[0][gbase + pos] = key;
      sc.val[0][gbase + pos] = v;
    }
    ++pos;
  };
#pragma unroll
  for (uint32_t k = 0; k < kAhead; ++k)
    if (k < c) put(ak[k], make_uint2(aa[k], ub));
  for (uint32_t k = kAhead; k < c; ++k) put(sc.node_slot[(size_t)k * L + t], make_uint2(sc.node_av[(size_t)k * L + t], ub));
  if (staged) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < total; i += 256) {
      sc.key[0][gbase + i] = s_key[i];
      sc.val[0][gbase + i] = s_val[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, one digit of at most kDigitBits bits per pass: tile histograms, a
// scan of the (digit, tile) counts, and a scatter that ranks each key within
// its tile by wave-level digit matching (one ballot per digit bit) so equal
// digits keep their order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t n_tiles_of(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }

// Workgroups are dispatched round-robin over the 8 XCDs by workgroup index.
// Work item i (i = blockIdx.x + k * gridDim.x, gridDim.x a multiple of 8)
// names tile (i % 8) * per + i / 8, so every XCD holds one contiguous range of
// tiles: a digit's output segments from neighbouring tiles then meet in that
// XCD's L2 and leave it as whole lines.  Returns nt for items past the end.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t i, uint32_t nt) {
  const uint32_t per = (nt + kXcds - 1) / kXcds;
  const uint32_t j = i / kXcds;
  const uint32_t tile = (i % kXcds) * per + j;
  return (j < per && tile < nt) ? tile : nt;
}
__device__ __forceinline__ uint32_t xcd_items(uint32_t nt) { return (nt + kXcds - 1) / kXcds * kXcds; }

// Lanes of the wave whose digit equals this lane's, among `valid` lanes.
template <uint32_t BITS>
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < (int)BITS; ++b) {
    const uint64_t bal = __ballot((d >> b) & 1u);
    peers &= ((d >> b) & 1u) ? bal : ~bal;
  }
  return peers;
}

template <uint32_t BITS>
__global__ __launch_bounds__(kSortWG) void k_rs_hist(const uint32_t* __restrict__ keys,
                                                     const LaunchCounters* __restrict__ cnt, uint32_t shift,
                                                     uint32_t* __restrict__ hist, uint32_t max_tiles) {
  constexpr uint32_t Bins = 1u << BITS;
  __shared__ uint32_t s_h[Bins];
  const uint32_t n = cnt->n_sort;
  const uint32_t nt = n_tiles_of(n);
  const int tx = threadIdx.x, lane = tx & 63;
  const uint64_t lt = (1ull << lane) - 1ull;
  const uint32_t items = xcd_items(nt);
  for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
    const uint32_t tile = xcd_tile(it, nt);
    if (tile >= nt) continue;  // workgroup-uniform
    uint32_t kd[kSortItems];  // the tile's keys, loaded before the counts are cleared
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
      const uint32_t idx = tile * kSortTile + (uint32_t)(i * kSortWG + tx);
      kd[i] = idx < n ? keys[idx] : 0u;
    }
    for (uint32_t b = tx; b < Bins; b += kSortWG) s_h[b] = 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
      const uint32_t idx = tile * kSortTile + (uint32_t)(i * kSortWG + tx);
      const bool valid = idx < n;
      const uint32_t d = valid ? (kd[i] >> shift) & (Bins - 1u) : 0u;
      const uint64_t peers = digit_peers<BITS>(d, valid);  // one LDS add per distinct digit of the wave
      if (valid && (peers & lt) == 0) atomicAdd(&s_h[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    for (uint32_t b = tx; b < Bins; b += kSortWG) hist[(size_t)b * max_tiles + tile] = s_h[b];
    __syncthreads();
  }
}

// One workgroup per digit: exclusive prefix of the digit's tile counts in
// place (coalesced chunks of 1024 tiles) and the digit's total.
__global__ __launch_bounds__(256) void k_rs_scan(uint32_t* __restrict__ hist, uint32_t* __restrict__ bin_total,
                                                 const LaunchCounters* __restrict__ cnt, uint32_t max_tiles) {
  __shared__ uint32_t s_w[4];
  const uint32_t nt = n_tiles_of(cnt->n_sort);
  uint32_t* row = hist + (size_t)blockIdx.x * max_tiles;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < nt; c0 += 1024) {
    uint32_t v[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t idx = c0 + threadIdx.x * 4 + (uint32_t)k;
      v[k] = idx < nt ? row[idx] : 0u;
      sum += v[k];
    }
    uint32_t tot = 0;
    uint32_t run = carry + block_excl_scan_256(sum, s_w, &tot);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t idx = c0 + threadIdx.x * 4 + (uint32_t)k;
      if (idx < nt) row[idx] = run;
      run += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) bin_total[blockIdx.x] = carry;
}

template <uint32_t BITS>
__global__ __launch_bounds__(kSortWG) void k_rs_scatter(const uint32_t* __restrict__ kin, const uint2* __restrict__ vin,
                                                        uint32_t* __restrict__ kout, uint2* __restrict__ vout,
                                                        const LaunchCounters* __restrict__ cnt, uint32_t shift,
                                                        const uint32_t* __restrict__ hist,
                                                        const uint32_t* __restrict__ bin_total, uint32_t max_tiles) {
  constexpr uint32_t Bins = 1u << BITS;
  static_assert(kSortWG == 256 && Bins % kSortWG == 0, "whole digits per thread");
  constexpr int kWaves = kSortWG / 64;
  constexpr int kPerThread = (int)(Bins / kSortWG);  // digits owned by each thread
  __shared__ uint32_t s_wc[kWaves][Bins];  // per-wave digit counts, then per-wave bases
  __shared__ uint32_t s_ls[Bins];          // tile-local start of each digit
  __shared__ uint32_t s_adj[Bins];         // global position of (digit, tile) minus its tile-local start
  __shared__ uint32_t s_k[kSortTile];      // the tile in digit order
  __shared__ uint2 s_v[kSortTile];
  __shared__ uint32_t s_w[4];
  const uint32_t n = cnt->n_sort;
  const uint32_t nt = n_tiles_of(n);
  const int tx = threadIdx.x, lane = tx & 63, w = tx >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  // first output position of each digit: exclusive scan of the digit totals
  // (thread tx owns digits tx*kPerThread .. +kPerThread-1), formed at the
  // workgroup's first tile
  uint32_t digit_base[kPerThread];
  bool based = false;
  const uint32_t items = xcd_items(nt);
  for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
    const uint32_t tile = xcd_tile(it, nt);
    if (tile >= nt) continue;  // workgroup-uniform
    // the tile's loads first (keys, values, this thread's digits' tile
    // offsets), so their latency overlaps the scans and the ranking set-up;
    // wave w holds keys [wbase, wbase + 64 * kSortItems) of the tile, item-major
    uint32_t kk[kSortItems], off[kSortItems];
    uint2 vv[kSortItems];
    const uint32_t wbase = tile * kSortTile + (uint32_t)(w * 64 * kSortItems);
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
      const uint32_t idx = wbase + (uint32_t)(i * 64 + lane);
      const bool valid = idx < n;
      kk[i] = valid ? kin[idx] : 0u;
      vv[i] = valid ? vin[idx] : make_uint2(0u, 0u);
    }
    uint32_t toff[kPerThread];
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) toff[q] = hist[(size_t)(tx * kPerThread + q) * max_tiles + tile];
    if (!based) {  // workgroup-uniform
      uint32_t sum = 0;
#pragma unroll
      for (int q = 0; q < kPerThread; ++q) {
        digit_base[q] = sum;
        sum += bin_total[tx * kPerThread + q];
      }
      const uint32_t b0 = block_excl_scan_256(sum, s_w, nullptr);
#pragma unroll
      for (int q = 0; q < kPerThread; ++q) digit_base[q] += b0;
      based = true;
    }
#pragma unroll
    for (int k = 0; k < kWaves; ++k)
#pragma unroll
      for (int q = 0; q < kPerThread; ++q) s_wc[k][tx * kPerThread + q] = 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
      const uint32_t idx = wbase + (uint32_t)(i * 64 + lane);
      const bool valid = idx < n;
      const uint32_t d = (kk[i] >> shift) & (Bins - 1u);
      const uint64_t peers = digit_peers<BITS>(d, valid);
      // all lanes read the wave's count before the group's first lane bumps it
      const uint32_t before = s_wc[w][d];
      off[i] = before + (uint32_t)__popcll(peers & lt);
      __builtin_amdgcn_wave_barrier();
      if (valid && (peers & lt) == 0) s_wc[w][d] = before + (uint32_t)__popcll(peers);
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: the waves' bases within the digit, the digit's count in
    // the tile, then the tile-local digit starts (a scan over the digits)
    uint32_t lstart[kPerThread];
    uint32_t tsum = 0;
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      const uint32_t dd = (uint32_t)(tx * kPerThread + q);
      uint32_t run = 0;
#pragma unroll
      for (int k = 0; k < kWaves; ++k) {
        const uint32_t c = s_wc[k][dd];
        s_wc[k][dd] = run;
        run += c;
      }
      lstart[q] = tsum;
      tsum += run;
    }
    const uint32_t lb = block_excl_scan_256(tsum, s_w, nullptr);
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      const uint32_t dd = (uint32_t)(tx * kPerThread + q);
      s_ls[dd] = lb + lstart[q];
      s_adj[dd] = digit_base[q] + toff[q] - (lb + lstart[q]);  // >= 0: earlier digits' keys all precede
    }
    __syncthreads();
    // the tile in digit order in LDS (equal digits keep their order), then
    // written out with consecutive threads on consecutive positions of each
    // digit's output segment
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
      const uint32_t idx = wbase + (uint32_t)(i * 64 + lane);
      if (idx < n) {
        const uint32_t d = (kk[i] >> shift) & (Bins - 1u);
        const uint32_t lp = s_ls[d] + s_wc[w][d] + off[i];
        s_k[lp] = kk[i];
        s_v[lp] = vv[i];
      }
    }
    __syncthreads();
    const uint32_t tcount = min(kSortTile, n - tile * kSortTile);
    for (uint32_t j = (uint32_t)tx; j < tcount; j += kSortWG) {
      const uint32_t key = s_k[j];
      const uint32_t pos = s_adj[(key >> shift) & (Bins - 1u)] + j;
      kout[pos] = key;
      vout[pos] = s_v[j];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Owner-computes table update (cfr_gpu.cu:784-837 for the contributions,
// :1018-1022 and :318-351 for clamp and re-matching).  Every run of equal
// keys is one slot of one tenant with exactly one owner (a lane for a short
// run, the wave of k_reduce_runs for a long one) holding the actions' FP64
// sums.  sigma comes from the pre-launch regret row, which nothing has
// modified yet: only the owner writes a slot.
// ---------------------------------------------------------------------------
// Finds the runs: one wave per 64 sorted records, head flags by ballot.  A
// run that ends inside its wave or among the next kShortRun records, with at
// most kShortRun records, is listed in the wave's region for k_reduce_rows;
// the longer ones (at most 4 per wave) in the wave's region for
// reduce_cand.  Plain stores, no atomics: every region has one writer.
// The heads of a wave's deferred runs lie more than kShortRun apart (all but
// the last are closed and long), so a wave defers at most 63 / (kShortRun + 1)
// + 1 of them, and its list holds 4; the look-ahead is one ballot.
static_assert(63 / (kShortRun + 1) + 1 <= 4 && kShortRun < 64, "a wave's deferred runs fit its list of 4");
__global__ __launch_bounds__(256) void k_runs(DevScratch sc, const uint32_t* __restrict__ keys, DevSeg* advance,
                                             const DevGame* __restrict__ tenants, uint32_t n_segs) {
  const uint32_t n = sc.cnt->n_rec;
  const int lane = threadIdx.x & 63;
  const uint32_t stride = gridDim.x * blockDim.x;
  // a captured training launch moves its segments on to the next launch here
  // (every kernel that reads them has run)
  if (advance)
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_segs; i += stride) {
      DevSeg sg = advance[i];
      advance_segment(sg, tenants[sg.tenant]);
      advance[i] = sg;
    }
  const uint64_t lt = (1ull << lane) - 1ull;
  for (uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; base < n; base += stride) {
    const uint32_t w = base >> 6;
    const uint32_t i = base + (uint32_t)lane;
    const bool valid = i < n;
    const uint32_t key = valid ? keys[i] : 0u;
    const bool head = valid && (i == 0 || keys[i - 1] != key);
    const uint64_t heads = __ballot(head);
    // run ends among the next kShortRun records (a head there, or the end of
    // the records), so a short run crossing into the next region is closed here
    const uint32_t xi = base + 64u + (uint32_t)lane;
    const uint32_t xkey = (lane < (int)kShortRun && xi < n) ? keys[xi] : 0u;
    // the record before extension lane l: lane 63's record for l = 0, else extension lane l - 1
    const uint32_t xprev = __shfl(lane == 63 ? key : xkey, lane == 0 ? 63 : lane - 1, 64);
    const uint64_t xends = __ballot(lane < (int)kShortRun && xi <= n && (xi == n || xkey != xprev)) &
                           ((1ull << kShortRun) - 1ull);
    const uint64_t later = heads & ~((2ull << lane) - 1ull);  // heads of higher lanes
    const uint32_t e = later ? base + (uint32_t)__builtin_ctzll(later)
                             : (xends ? base + 64u + (uint32_t)__builtin_ctzll(xends) : min(base + 64u, n));
    const bool closed = later != 0 || xends != 0 || e == n;
    const bool is_short = head && closed && e - i <= kShortRun;
    const bool defer = head && !is_short;
    const uint64_t sm = __ballot(is_short);
    const uint64_t dm = __ballot(defer);
    if (is_short) sc.runs[(size_t)w * 64 + (uint32_t)__popcll(sm & lt)] = make_uint4(i, e - i, key, 0u);
    if (defer) sc.cand[(size_t)w * 4 + (uint32_t)__popcll(dm & lt)] = make_uint2(i, closed ? e : 0u);
    if (lane == 0) {
      sc.n_runs[w] = (uint32_t)__popcll(sm);
      sc.n_cand[w] = (uint32_t)__popcll(dm);
      sc.head0[w] = heads ? (uint8_t)__builtin_ctzll(heads) : (uint8_t)64;
      if (heads) sc.head1[w >> 6] = 0;
    }
  }
}

// The first of the 64 bytes at base + a (a a multiple of 64) from index
// `from` on that is below 64, an index at or past `lim` counting as one:
// its offset from a (64 if none) and, in *val, the byte.
__device__ __forceinline__ uint32_t first_head64(const uint8_t* __restrict__ base, uint32_t a, uint32_t from,
                                                 uint32_t lim, uint32_t* val) {
  const uint4* p = reinterpret_cast<const uint4*>(base + a);
  uint32_t wd[16];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint4 v = p[q];
    wd[4 * q] = v.x;
    wd[4 * q + 1] = v.y;
    wd[4 * q + 2] = v.z;
    wd[4 * q + 3] = v.w;
  }
  uint32_t hit = 64;
  *val = 0;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const uint32_t rb = a + 4u * (uint32_t)q;  // index of the word's byte 0
    uint32_t hm = ~wd[q] & 0x40404040u;       // bytes 0..63 (64 has bit 6 set)
    hm &= from > rb ? (from - rb >= 4u ? 0u : 0xFFFFFFFFu << (8u * (from - rb))) : 0xFFFFFFFFu;
    hm |= lim > rb ? (lim - rb >= 4u ? 0u : 0x40404040u & (0xFFFFFFFFu << (8u * (lim - rb)))) : 0x40404040u;
    if (hit == 64u && hm != 0u) {
      const uint32_t k = (uint32_t)__builtin_ctz(hm) >> 3;
      hit = 4u * (uint32_t)q + k;
      *val = (wd[q] >> (8u * k)) & 0xFFu;
    }
  }
  return hit;
}

// One thread per wave region's deferred runs (k_runs defers only runs of more
// than kShortRun records, and passes on the end of those that end within the
// region or the next kShortRun records).  Any other deferred run ends at the
// first run head of a later region, or at the end of the records: found in
// the region's group of 64 regions (head0, one byte each, one 64-byte read),
// else through the first later group that holds a head (head1, one byte per
// group) and that group's regions.  Each run is cut into kPiece-record
// pieces.  A wave reserves the list space of all its long runs and pieces
// with one 64-bit atomic (the long-run and piece counters share a word); a
// lane writes the pieces of its runs of up to 4 pieces, the wave those of
// longer runs.
// Runs as the leading `nblk` workgroups of k_reduce_rows (the two depend only
// on k_runs), `blk` being the workgroup's index among them.
__device__ __forceinline__ void reduce_cand(DevScratch sc, const uint32_t* __restrict__ keys, uint32_t blk,
                                            uint32_t nblk) {
  const uint32_t n = sc.cnt->n_rec;
  const uint32_t n_waves = (n + 63) / 64;
  const int lane = threadIdx.x & 63;
  const uint32_t stride = nblk * blockDim.x;
  for (uint32_t w0 = (blk * blockDim.x + threadIdx.x) & ~63u; w0 < n_waves; w0 += stride) {
    const uint32_t w = w0 + (uint32_t)lane;
    const uint32_t nc = w < n_waves ? sc.n_cand[w] : 0u;
    // the region's 4 candidate slots, loaded beside their count (slots past
    // the count are not used)
    uint2 cds[4];
    if (w < n_waves) {
      const uint4* cp = reinterpret_cast<const uint4*>(sc.cand + (size_t)w * 4);
      const uint4 c01 = cp[0], c23 = cp[1];
      cds[0] = make_uint2(c01.x, c01.y);
      cds[1] = make_uint2(c01.z, c01.w);
      cds[2] = make_uint2(c23.x, c23.y);
      cds[3] = make_uint2(c23.z, c23.w);
    }
    uint32_t lstart[4], llen[4], lkey[4];  // long runs of the region (llen 0: none)
    uint32_t nl = 0, np = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      llen[c] = 0;
      lstart[c] = 0;
      lkey[c] = 0;
      if ((uint32_t)c >= nc) continue;
      const uint32_t i = cds[c].x;
      uint32_t e = cds[c].y;
      lkey[c] = keys[i];  // in flight during the search
      if (e == 0) {
        uint32_t v;
        const uint32_t a = (w + 1) & ~63u;
        const uint32_t h = first_head64(sc.head0, a, w + 1, n_waves, &v);
        if (h < 64u) {
          e = a + h >= n_waves ? n : (a + h) * 64u + v;
        } else {
          const uint32_t n_groups = (n_waves + 63) / 64;
          uint32_t g = (a >> 6) + 1, ga, gh;
          for (;;) {  // each pass moves g on by at least one; groups from n_groups on are hits
            ga = g & ~63u;
            gh = first_head64(sc.head1, ga, g, n_groups, &v);
            if (gh < 64u) break;
            g = ga + 64u;
          }
          const uint32_t grp = ga + gh;
          if (grp >= n_groups) {
            e = n;
          } else {  // the group holds a head (or regions past the records)
            const uint32_t h2 = first_head64(sc.head0, grp * 64u, grp * 64u, n_waves, &v);
            e = grp * 64u + h2 >= n_waves ? n : (grp * 64u + h2) * 64u + v;
          }
        }
      }
      lstart[c] = i;
      llen[c] = e - i;
      ++nl;
      np += (e - i + kPiece - 1) / kPiece;
    }
    // the wave's exclusive prefix of (pieces << 32 | long runs), one atomic
    const uint64_t mine = ((uint64_t)np << 32) | nl;
    uint64_t incl = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t o = __shfl_up(incl, d, 64);
      if (lane >= d) incl += o;
    }
    const uint64_t total = __shfl(incl, 63, 64);
    uint64_t base = 0;
    if (lane == 0 && total) base = atomicAdd(reinterpret_cast<unsigned long long*>(&sc.cnt->n_long), total);
    base = __shfl(base, 0, 64);
    const uint64_t excl = base + incl - mine;
    uint32_t r = (uint32_t)excl, p0 = (uint32_t)(excl >> 32);
    uint32_t rc[4], pc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      rc[c] = r;
      pc[c] = p0;
      if (llen[c] == 0) continue;
      const uint32_t i = lstart[c], end = i + llen[c], npc = (llen[c] + kPiece - 1) / kPiece;
      sc.long_runs[r] = make_uint4(i, llen[c], p0, npc);
      if (npc <= 4)
        for (uint32_t j = 0; j < npc; ++j)
          sc.piece[p0 + j] = make_uint4(i + j * kPiece, min(end, i + (j + 1) * kPiece), lkey[c], r);
      ++r;
      p0 += npc;
    }
    // the pieces of runs of more than 4 pieces, written by the whole wave
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      uint64_t m = __ballot(llen[c] > 4 * kPiece);
      while (m) {
        const int src = (int)__builtin_ctzll(m);
        m &= m - 1;
        const uint32_t i = __shfl(lstart[c], src, 64), len = __shfl(llen[c], src, 64);
        const uint32_t p = __shfl(pc[c], src, 64), rr = __shfl(rc[c], src, 64), key = __shfl(lkey[c], src, 64);
        const uint32_t npc = (len + kPiece - 1) / kPiece;
        for (uint32_t j = (uint32_t)lane; j < npc; j += 64)
          sc.piece[p + j] = make_uint4(i + j * kPiece, min(i + len, i + (j + 1) * kPiece), key, rr);
      }
    }
  }
}

// Lane a of a ROW-lane group: R <- R + dR (CFR+ clamp), S <- S + dS, and the
// slot's BF16 strategy re-matched from the new row (sum in action order).
template <int ROW>
__device__ __forceinline__ void apply_lane(const SegCtx& c, size_t slot, int a, int group_lane0, float r_a,
                                           float s_a, double dr, double ds) {
  float nr = (float)((double)r_a + dr);
  if (c.cfr_plus) nr = fmaxf(0.0f, nr);
  const float ns = (float)((double)s_a + ds);
  float sum = 0.0f;
#pragma unroll
  for (int b = 0; b < ROW; ++b) sum += fmaxf(0.0f, __shfl(nr, group_lane0 + b, 64));
  const float x = (sum > 1e-7f) ? fmaxf(0.0f, nr) / sum : 0.0f;
  gbl(c.R)[slot * 2 * ROW + a] = nr;
  gbl(c.S)[slot * 2 * ROW + a] = ns;
  gbl(c.sig)[slot * ROW + a] = f32_to_bf16(x);
}

__device__ __forceinline__ SegCtx ctx_of_key(const SegCtx* __restrict__ sctx, uint32_t key, uint32_t slot_bits,
                                             uint32_t* slot) {
  *slot = key & (uint32_t)((1ull << slot_bits) - 1ull);
  return sctx[(uint32_t)((uint64_t)key >> slot_bits)];
}

// Short runs, one lane each.  A wave takes two neighbouring wave regions' run
// lists (at most 64 runs each, ~48 together) and lane j the pair's j-th run;
// the records those runs can reach are staged in LDS with coalesced loads.
// The lane holds the slot's regret and strategy-sum rows, sums every action's
// contributions in FP64 in record order, and stores the updated rows and the
// re-matched BF16 strategy whole.  Per record and legal action a:
//   dR_a += a == sampled ? u / q - u : 0 - u,   q = max(sigma_sampled, 0.05)
//   dS_a += tw * sigma_a
// with sigma from the pre-launch regret row and the record's legal mask
// (cfr_gpu.cu:784-801, :828).  dS_a is the same for every record of a mask,
// so it is added once per stretch of equal masks as count * (tw * sigma_a):
// exact in FP64, and equal to the record-by-record sum for a run of one mask
// (the rare earlier stretches of a run wait in LDS).
// A lane per run keeps a wave's row gathers in flight together; row accesses
// are latency-bound, not bandwidth-bound.  8-wide rows are held in registers
// through the record loop (96 registers, 5 waves per SIMD; variants with 6
// waves measured the same); 16-wide rows are re-read (below).
// The first `cand_blocks` workgroups find the long runs' ends and pieces
// instead (reduce_cand): that latency-bound search overlaps the short runs'
// updates rather than taking a launch of its own.
template <int ROW, bool ES>
__global__ __launch_bounds__(256) void k_reduce_rows(LaunchCfg lc, DevScratch sc, const uint32_t* __restrict__ keys,
                                                     const uint2* __restrict__ vals, uint32_t cand_blocks) {
  if (blockIdx.x < cand_blocks) {
    reduce_cand(sc, keys, blockIdx.x, cand_blocks);
    return;
  }
  const uint32_t blk = blockIdx.x - cand_blocks;
  constexpr int kWaves = 256 / 64;
  // runs of a pair of regions start among its 128 records and hold at most kShortRun
  constexpr uint32_t kSpan = 2 * 64 + kShortRun;
  __shared__ uint2 s_v[kWaves][kSpan];
  __shared__ double s_ds[ROW][256];
  const uint32_t n = sc.cnt->n_rec;
  const uint32_t n_regions = (n + 63) / 64;
  const uint32_t n_pairs = (n_regions + 1) / 2;
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  const bool one_seg = lc.n_segs == 1;
  SegCtx c0{};
  if (one_seg) c0 = sc.seg_ctx[0];
  const uint32_t slot_mask = (uint32_t)((1ull << lc.slot_bits) - 1ull);
  const uint32_t wstride = (gridDim.x - cand_blocks) * kWaves;
  for (uint32_t pw = blk * kWaves + wv; pw < n_pairs; pw += wstride) {
    const uint32_t r0 = 2 * pw, base = r0 * 64u;
    for (uint32_t k = lane; k < kSpan; k += 64) s_v[wv][k] = base + k < n ? vals[base + k] : make_uint2(0u, 0u);
    const uint32_t nr0 = sc.n_runs[r0];
    const uint32_t nr1 = r0 + 1 < n_regions ? sc.n_runs[r0 + 1] : 0u;
    wave_lds_sync();
    for (uint32_t j = lane; j < nr0 + nr1; j += 64) {
      const uint4 run = sc.runs[j < nr0 ? (size_t)r0 * 64 + j : (size_t)(r0 + 1) * 64 + (j - nr0)];
      const uint32_t slot = run.z & slot_mask;
      const SegCtx c = one_seg ? c0 : sc.seg_ctx[(uint32_t)((uint64_t)run.z >> lc.slot_bits)];
      if constexpr (ROW <= 8) {
        const size_t rs_off = (size_t)slot * 2 * ROW;  // interleaved regret / strategy-sum row
        const int A = c.num_actions;
        float rr[ROW];
        load_f32_row<ROW>(c.R + rs_off, rr);
        double dr[ROW];
#pragma unroll
        for (int b = 0; b < ROW; ++b) dr[b] = 0.0;
        // the current mask's sigma per action (a legal action's share of the
        // positive regrets, or the uniform share), formed once per mask: q is
        // max(sigma, 0.05) and a stretch's dS uses sigma itself
        uint32_t cur_vm = 0, cnt = 0;
        float sv[ROW];
#pragma unroll
        for (int b = 0; b < ROW; ++b) sv[b] = 0.0f;
        // dS of a mask stretch: count * (tw * sigma_a) for the legal actions
        auto stretch_ds = [&](int b) -> double {
          const bool legal = b < A && (cur_vm >> b & 1u);
          return legal ? (double)cnt * (double)(c.tw * sv[b]) : 0.0;
        };
        bool spilled = false;  // dS of earlier stretches (a mask change inside the run) is in s_ds
        for (uint32_t k = 0; k < run.y; ++k) {
          const uint2 v = s_v[wv][run.x - base + k];
          const uint32_t act = v.x & 0xFFu, vm = (v.x >> 8) & kLegalBits;
          const bool es = ES && (v.x & kEsRecord) != 0;
          const float u = __uint_as_float(v.y);
          if (vm != cur_vm) {
            if (cnt) {
#pragma unroll
              for (int b = 0; b < ROW; ++b) s_ds[b][threadIdx.x] = (spilled ? s_ds[b][threadIdx.x] : 0.0) + stretch_ds(b);
              spilled = true;
            }
            cnt = 0;
            cur_vm = vm;
            float t = 0.0f;
#pragma unroll
            for (int b = 0; b < ROW; ++b) t += (b < A && (vm >> b & 1u)) ? fmaxf(0.0f, rr[b]) : 0.0f;
            const float uni = 1.0f / (float)__popc(vm);
            const bool pos = t > 1e-7f;
#pragma unroll
            for (int b = 0; b < ROW; ++b) sv[b] = pos ? fmaxf(0.0f, rr[b]) / t : uni;
          }
          ++cnt;  // every record is one node
          if (es) {
            // external sampling: the node's row of regret deltas
            float rdv[ROW];
            load_f32_row<ROW>(sc.es_rd + (size_t)v.y * ROW, rdv);
#pragma unroll
            for (int b = 0; b < ROW; ++b)
              if (b < A && (vm >> b & 1u)) dr[b] += (double)rdv[b];
          } else {
            float sa = sv[0];
#pragma unroll
            for (int b = 1; b < ROW; ++b) sa = ((uint32_t)b == act) ? sv[b] : sa;
            const float q = fmaxf(sa, 0.05f);
            // outcome sampling: u / q - u for the sampled action, -u for the
            // other legal ones (each rounded to FP32, then widened once)
            const double rd_sampled = (double)(u / q - u);
            const double rd_other = (double)(0.0f - u);
#pragma unroll
            for (int b = 0; b < ROW; ++b)
              if (b < A && (vm >> b & 1u)) dr[b] += (uint32_t)b == act ? rd_sampled : rd_other;
          }
        }
        float ss[ROW];
        load_f32_row<ROW>(c.S + rs_off, ss);
        double ds[ROW];
#pragma unroll
        for (int b = 0; b < ROW; ++b) ds[b] = (spilled ? s_ds[b][threadIdx.x] : 0.0) + stretch_ds(b);
        // R <- R + dR (CFR+ clamp), S <- S + dS, sigma re-matched from the new row
        float nr[ROW], ns[ROW];
        float sum = 0.0f;
#pragma unroll
        for (int b = 0; b < ROW; ++b) {
          nr[b] = (float)((double)rr[b] + dr[b]);
          if (c.cfr_plus) nr[b] = fmaxf(0.0f, nr[b]);
          ns[b] = (float)((double)ss[b] + ds[b]);
          sum += fmaxf(0.0f, nr[b]);
        }
        uint32_t sg[ROW / 2];
#pragma unroll
        for (int b = 0; b < ROW; b += 2) {
          const float x0 = (sum > 1e-7f) ? fmaxf(0.0f, nr[b]) / sum : 0.0f;
          const float x1 = (sum > 1e-7f) ? fmaxf(0.0f, nr[b + 1]) / sum : 0.0f;
          sg[b / 2] = (uint32_t)f32_to_bf16(x0) | ((uint32_t)f32_to_bf16(x1) << 16);
        }
        store_f32_row<ROW>(c.R + rs_off, nr);
        store_f32_row<ROW>(c.S + rs_off, ns);
        store_bf16_pairs<ROW>(c.sig + (size_t)slot * ROW, sg);
      } else {
        const size_t rs_off = (size_t)slot * 2 * ROW;  // interleaved regret / strategy-sum row
        const int A = c.num_actions;
        double dr[ROW];
#pragma unroll
        for (int b = 0; b < ROW; ++b) dr[b] = 0.0;
        // the current mask's sigma per action, formed once per mask (q is
        // max(sigma, 0.05)).  The 16-wide regret row is read again where it
        // is needed (at a mask change and for the apply) rather than held
        // through the record loop: 176 -> 114 registers, 2 -> 4 waves per SIMD.
        uint32_t cur_vm = 0, cnt = 0;
        float sv[ROW];
#pragma unroll
        for (int b = 0; b < ROW; ++b) sv[b] = 0.0f;
        // dS of a mask stretch: count * (tw * sigma_a) for the legal actions
        auto stretch_ds = [&](int b) -> double {
          const bool legal = b < A && (cur_vm >> b & 1u);
          return legal ? (double)cnt * (double)(c.tw * sv[b]) : 0.0;
        };
        bool spilled = false;  // dS of earlier stretches (a mask change inside the run) is in s_ds
        for (uint32_t k = 0; k < run.y; ++k) {
          const uint2 v = s_v[wv][run.x - base + k];
          const uint32_t act = v.x & 0xFFu, vm = (v.x >> 8) & kLegalBits;
          const bool es = ES && (v.x & kEsRecord) != 0;
          const float u = __uint_as_float(v.y);
          if (vm != cur_vm) {
            float rr[ROW];
            load_f32_row<ROW>(c.R + rs_off, rr);
            if (cnt) {
#pragma unroll
              for (int b = 0; b < ROW; ++b)
                s_ds[b][threadIdx.x] = (spilled ? s_ds[b][threadIdx.x] : 0.0) + stretch_ds(b);
              spilled = true;
            }
            cnt = 0;
            cur_vm = vm;
            float t = 0.0f;
#pragma unroll
            for (int b = 0; b < ROW; ++b) t += (b < A && (vm >> b & 1u)) ? fmaxf(0.0f, rr[b]) : 0.0f;
            const float uni = 1.0f / (float)__popc(vm);
            const bool pos = t > 1e-7f;
#pragma unroll
            for (int b = 0; b < ROW; ++b) sv[b] = pos ? fmaxf(0.0f, rr[b]) / t : uni;
          }
          ++cnt;  // every record is one node
          if (es) {
            // external sampling: the node's row of regret deltas
            float rdv[ROW];
            load_f32_row<ROW>(sc.es_rd + (size_t)v.y * ROW, rdv);
#pragma unroll
            for (int b = 0; b < ROW; ++b)
              if (b < A && (vm >> b & 1u)) dr[b] += (double)rdv[b];
          } else {
            float sa = sv[0];
#pragma unroll
            for (int b = 1; b < ROW; ++b) sa = ((uint32_t)b == act) ? sv[b] : sa;
            const float q = fmaxf(sa, 0.05f);
            // outcome sampling: u / q - u for the sampled action, -u for the
            // other legal ones (each rounded to FP32, then widened once)
            const double rd_sampled = (double)(u / q - u);
            const double rd_other = (double)(0.0f - u);
#pragma unroll
            for (int b = 0; b < ROW; ++b)
              if (b < A && (vm >> b & 1u)) dr[b] += (uint32_t)b == act ? rd_sampled : rd_other;
          }
        }
        asm volatile("" ::: "memory");  // the rows below are loaded here, not kept from the loop
        // R <- R + dR (CFR+ clamp), S <- S + dS four actions at a time, loads of
        // a chunk after the previous chunk's stores (only the new regrets stay
        // live, for the re-match); sigma re-matched from the new row (sum in
        // action order)
        float rr[ROW], ss[ROW];
        float nr[ROW];
        float sum = 0.0f;
        const auto qr = gbl(reinterpret_cast<f32x4*>(c.R + rs_off));
        const auto qs = gbl(reinterpret_cast<f32x4*>(c.S + rs_off));
#pragma unroll
        for (int q4 = 0; q4 < ROW / 4; ++q4) {
          const f32x4 r4 = qr[q4], s4 = qs[q4];
          rr[4 * q4] = r4.x;
          rr[4 * q4 + 1] = r4.y;
          rr[4 * q4 + 2] = r4.z;
          rr[4 * q4 + 3] = r4.w;
          ss[4 * q4] = s4.x;
          ss[4 * q4 + 1] = s4.y;
          ss[4 * q4 + 2] = s4.z;
          ss[4 * q4 + 3] = s4.w;
          float ns[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int b = 4 * q4 + j;
            const double ds = (spilled ? s_ds[b][threadIdx.x] : 0.0) + stretch_ds(b);
            float x = (float)((double)rr[b] + dr[b]);
            if (c.cfr_plus) x = fmaxf(0.0f, x);
            nr[b] = x;
            ns[j] = (float)((double)ss[b] + ds);
            sum += fmaxf(0.0f, x);
          }
          f32x4 o;
          o.x = nr[4 * q4];
          o.y = nr[4 * q4 + 1];
          o.z = nr[4 * q4 + 2];
          o.w = nr[4 * q4 + 3];
          qr[q4] = o;
          f32x4 o2;
          o2.x = ns[0];
          o2.y = ns[1];
          o2.z = ns[2];
          o2.w = ns[3];
          qs[q4] = o2;
        }
        uint32_t sg[ROW / 2];
#pragma unroll
        for (int b = 0; b < ROW; b += 2) {
          const float x0 = (sum > 1e-7f) ? fmaxf(0.0f, nr[b]) / sum : 0.0f;
          const float x1 = (sum > 1e-7f) ? fmaxf(0.0f, nr[b + 1]) / sum : 0.0f;
          sg[b / 2] = (uint32_t)f32_to_bf16(x0) | ((uint32_t)f32_to_bf16(x1) << 16);
        }
        store_bf16_pairs<ROW>(c.sig + (size_t)slot * ROW, sg);
      }
    }
    wave_lds_sync();  // the next pair's staging overwrites s_v
  }
}

// One wave per piece of a long run (kPiece records), in two phases.  First a
// lane per record: the record's q and sampled-action regret u / q - u (one
// pair of divisions per record instead of one per record and action), staged
// in LDS with the record's action word and utility.  Then its 64 / ROW groups
// take the records in a fixed stride, lane a summing action a's contributions
// in FP64 (sigma cached per legal mask), and lane a of
// the first group adds the groups' sums in group order (shuffles); the
// piece's FP64 sums are stored.  A piece is defined from its run's first
// record, so the sums depend only on the run's records, not on where the run
// sits in the launch.
template <int ROW, bool ES>
__global__ __launch_bounds__(256) void k_reduce_piece(LaunchCfg lc, DevScratch sc, const uint32_t* __restrict__ keys,
                                                      const uint2* __restrict__ vals) {
  constexpr int kGroups = 64 / ROW;
  constexpr int kWaves = 256 / 64;
  __shared__ uint32_t s_av[kWaves][kPiece];
  __shared__ float s_rs[kWaves][kPiece], s_u[kWaves][kPiece];
  const uint32_t n_pieces = sc.cnt->n_pieces;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int a = lane & (ROW - 1), g = lane / ROW;
  const uint32_t wstride = gridDim.x * kWaves;
  for (uint32_t p = blockIdx.x * kWaves + (uint32_t)wv; p < n_pieces; p += wstride) {
    const uint4 pc = sc.piece[p];
    const uint32_t len = pc.y - pc.x;
    uint2 v[kPiece / 64];
#pragma unroll
    for (int b = 0; b < (int)(kPiece / 64); ++b) {
      const uint32_t k = (uint32_t)(lane + 64 * b);
      v[b] = k < len ? vals[pc.x + k] : make_uint2(0u, 0u);
    }
    uint32_t slot;
    const SegCtx c = ctx_of_key(sc.seg_ctx, pc.z, lc.slot_bits, &slot);
    const int A = c.num_actions;
    float rr[ROW];
    load_f32_row<ROW>(c.R + (size_t)slot * 2 * ROW, rr);
    // phase 1: one lane per record
    uint32_t m_vm = 0;
    float m_sum = 0.0f, m_uni = 0.0f;
#pragma unroll
    for (int b = 0; b < (int)(kPiece / 64); ++b) {
      const uint32_t act = v[b].x & 0xFFu, vm = (v[b].x >> 8) & kLegalBits;
      const bool es = ES && (v[b].x & kEsRecord) != 0;
      const float u = __uint_as_float(v[b].y);
      if (vm != m_vm) {
        m_vm = vm;
        float t = 0.0f;
#pragma unroll
        for (int q = 0; q < ROW; ++q) t += (q < A && (vm >> q & 1u)) ? fmaxf(0.0f, rr[q]) : 0.0f;
        m_sum = t;
        m_uni = 1.0f / (float)__popc(vm);
      }
      float r_act = rr[0];
#pragma unroll
      for (int q = 1; q < ROW; ++q) r_act = ((uint32_t)q == act) ? rr[q] : r_act;
      const float qa = fmaxf(m_sum > 1e-7f ? fmaxf(0.0f, r_act) / m_sum : m_uni, 0.05f);
      const uint32_t k = (uint32_t)(lane + 64 * b);
      s_av[wv][k] = v[b].x;
      // external sampling: the node's delta-row index in place of the regret
      s_rs[wv][k] = es ? __uint_as_float(v[b].y) : u / qa - u;
      s_u[wv][k] = es ? 0.0f : u;
    }
    wave_lds_sync();
    // phase 2: lane a of each group sums action a
    float ra = rr[0];
#pragma unroll
    for (int q = 1; q < ROW; ++q) ra = (q == a) ? rr[q] : ra;
    double dr = 0.0, ds = 0.0;
    uint32_t cached_vm = 0;
    float sig_a = 0.0f;
    bool legal = false;
    for (uint32_t k = (uint32_t)g; k < len; k += kGroups) {
      const uint32_t av = s_av[wv][k];
      const uint32_t act = av & 0xFFu, vm = (av >> 8) & kLegalBits;
      if (vm != cached_vm) {
        cached_vm = vm;
        float t = 0.0f;
#pragma unroll
        for (int q = 0; q < ROW; ++q) t += (q < A && (vm >> q & 1u)) ? fmaxf(0.0f, rr[q]) : 0.0f;
        const float uni = 1.0f / (float)__popc(vm);
        const bool pos = t > 1e-7f;
        legal = a < A && (vm >> a & 1u);
        sig_a = pos ? (legal ? fmaxf(0.0f, ra) : 0.0f) / t : (legal ? uni : 0.0f);
      }
      if (legal) {
        if (ES && (av & kEsRecord))
          dr += (double)sc.es_rd[(size_t)__float_as_uint(s_rs[wv][k]) * ROW + a];
        else
          dr += (double)((uint32_t)a == act ? s_rs[wv][k] : 0.0f - s_u[wv][k]);
        ds += (double)(c.tw * sig_a);  // every record is one node
      }
    }
    double tr = 0.0, ts = 0.0;
#pragma unroll
    for (int k = 0; k < kGroups; ++k) {
      tr += __shfl(dr, k * ROW + a, 64);
      ts += __shfl(ds, k * ROW + a, 64);
    }
    if (g == 0) {
      sc.partial[(size_t)p * 32 + a] = tr;
      sc.partial[(size_t)p * 32 + ROW + a] = ts;
    }
    wave_lds_sync();  // the next piece's staging overwrites the LDS arrays
  }
}

// One wave per long run: its 64 / ROW groups take the run's pieces in a fixed
// stride, the groups' sums are added in group order (shuffles), and the first
// group applies the update.
template <int ROW>
__global__ __launch_bounds__(256) void k_reduce_runs(LaunchCfg lc, DevScratch sc, const uint32_t* __restrict__ keys) {
  constexpr int kGroups = 64 / ROW;
  const uint32_t n_long = sc.cnt->n_long;
  const int lane = threadIdx.x & 63;
  const int a = lane & (ROW - 1), g = lane / ROW;
  const uint32_t wstride = gridDim.x * (blockDim.x / 64);
  for (uint32_t r = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); r < n_long; r += wstride) {
    const uint4 run = sc.long_runs[r];
    uint32_t slot;
    const SegCtx c = ctx_of_key(sc.seg_ctx, keys[run.x], lc.slot_bits, &slot);
    double dr = 0.0, ds = 0.0;
    for (uint32_t j = (uint32_t)g; j < run.w; j += kGroups) {
      dr += sc.partial[(size_t)(run.z + j) * 32 + a];
      ds += sc.partial[(size_t)(run.z + j) * 32 + ROW + a];
    }
    double tr = 0.0, ts = 0.0;
#pragma unroll
    for (int k = 0; k < kGroups; ++k) {
      tr += __shfl(dr, k * ROW + a, 64);
      ts += __shfl(ds, k * ROW + a, 64);
    }
    if (g == 0) {
      const float r_a = gbl(c.R)[(size_t)slot * 2 * ROW + a];
      const float s_a = gbl(c.S)[(size_t)slot * 2 * ROW + a];
      apply_lane<ROW>(c, slot, a, 0, r_a, s_a, tr, ts);
    }
  }
}

template <int ROW, int WG>
void walk_launch(hipStream_t s, const LaunchCfg& lc, const DevSeg* segs, const DevGame* tenants,
                 const DevScratch& sc, uint32_t tpl, mccfr_trace* trace) {
  uint32_t jstride = 64u, grid;
  if (tpl == 0 && lc.n_segs == 1) {
    // persistent: as many workgroups as the device keeps resident
    static const uint32_t resident = [] {
      int per_cu = 0, cus = 0, dev = 0;
      (void)hipGetDevice(&dev);
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_walk<ROW, WG, false>, WG, 0);
      return (uint32_t)(per_cu > 0 && cus > 0 ? per_cu * cus : 1);
    }();
    grid = std::min<uint32_t>(resident, (lc.lanes + WG - 1) / WG);
    jstride = grid * WG;
    tpl = (lc.lanes + jstride - 1) / jstride;
  } else {
    if (tpl == 0) tpl = 1;
    const uint32_t waves = lc.lanes / (64u * tpl);
    grid = (waves * 64u + WG - 1) / WG;
  }
  if (trace)
    hipLaunchKernelGGL((k_walk<ROW, WG, true>), dim3(grid), dim3(WG), 0, s, lc, segs, tenants, sc, tpl, jstride,
                       trace);
  else
    hipLaunchKernelGGL((k_walk<ROW, WG, false>), dim3(grid), dim3(WG), 0, s, lc, segs, tenants, sc, tpl, jstride,
                       trace);
}

template <int ROW>
void walk_dispatch(hipStream_t s, const LaunchCfg& lc, const DevSeg* segs, const DevGame* tenants,
                   const DevScratch& sc, WalkShape sh, mccfr_trace* trace) {
  switch (sh.wg) {
    case 64: walk_launch<ROW, 64>(s, lc, segs, tenants, sc, sh.tpl, trace); break;
    case 128: walk_launch<ROW, 128>(s, lc, segs, tenants, sc, sh.tpl, trace); break;
    case 512: walk_launch<ROW, 512>(s, lc, segs, tenants, sc, sh.tpl, trace); break;
    case 1024: walk_launch<ROW, 1024>(s, lc, segs, tenants, sc, sh.tpl, trace); break;
    default: walk_launch<ROW, 256>(s, lc, segs, tenants, sc, sh.tpl, trace); break;
  }
}

// ---------------------------------------------------------------------------
// External sampling by generations: the walk of one-segment launches
// (k_walk_es walks launches of several segments depth first; both explore
// the same tasks with the same streams, so their results are the same).
// Generation 0's tasks are the traversals' roots.  One lane walks a task as
// k_walk walks a traversal -- sampling the other seats' actions from the BF16
// snapshot with the task's own stream -- until a leaf, whose value goes to its
// parent node's child slot (or is the traversal's root value), or an
// update-player node.  Such a node gets an index (one atomic per wave on the
// generation's node count), stores its state as a 64-byte record and makes
// one task of the next generation per legal action, a 4-byte reference
// (node << 4 | action) in the other task buffer (one atomic per wave on the
// next generation's task count); the task applies its action to the node's
// state when it starts.  A node's child values collect in its row of
// DevScratch::es_rd; k_es_back then forms v(I) and the row of v(a) - v(I),
// generation by generation from the deepest, and passes v(I) to the parent.
// Every node is one record of the sort input, at its own index.  Past a
// buffer's capacity the launch is flagged (es_overflow) and the later
// generations do nothing: the launch updates no table.
// Node record: q0 = seat words 0..3 (chips behind | bet << 15 | street tag <<
// 30); q1 = seat words 4, 5, folded | all-in << 8 | queue << 16 | queue
// position << 24, history; q2 = pot, current bet | last full raise << 16,
// street | raises << 8 | depth << 16 | dealer << 24, traversal lane;
// q3 = parent node << 4 | action (kEsNoParent at a traversal's first node),
// its task's path code, slot, legal mask.
// ---------------------------------------------------------------------------

// Zeroes the generation counters and the launch's record counts.
__global__ __launch_bounds__(256) void k_es_gen_begin(DevScratch sc) {
  for (uint32_t i = threadIdx.x; i < (kEsGens + 1) * kEsRegions; i += blockDim.x) {
    sc.gen.task_cnt[i * kEsCtrStride] = 0;
    sc.gen.node_cnt[i * kEsCtrStride] = 0;
  }
  if (threadIdx.x == 0) {
    sc.cnt->n_es = 0;
    sc.cnt->n_es_valid = 0;
  }
}

__device__ __forceinline__ uint32_t& es_ctr(uint32_t* c, uint32_t g, uint32_t r) {
  return c[((size_t)g * kEsRegions + r) * kEsCtrStride];
}

// A generation's items (tasks or nodes) lie in kEsRegions regions with
// counts cnt(g, r).  Lane r of the calling wave holds region r's count
// (clamped to the region's capacity) and their inclusive prefix; item i of
// the generation is in the region whose prefix first exceeds i.
struct EsRegionScan {
  uint32_t count, incl, total;
};
__device__ __forceinline__ EsRegionScan es_region_scan(uint32_t* ctr, uint32_t g, uint32_t cap_r) {
  const int lane = threadIdx.x & 63;
  static_assert(kEsRegions == 64, "one region per lane");
  EsRegionScan rs;
  rs.count = min(es_ctr(ctr, g, (uint32_t)lane), cap_r);
  uint32_t incl = rs.count;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  rs.incl = incl;
  rs.total = uniform(__shfl(incl, 63, 64));
  return rs;
}
// Region of item i (< total) and its position within the region.
__device__ __forceinline__ uint32_t es_region_of(const EsRegionScan& rs, uint32_t i, uint32_t* pos) {
  uint32_t lo = 0, hi = 63;  // the first region whose inclusive prefix exceeds i
#pragma unroll
  for (int step = 0; step < 6; ++step) {
    const uint32_t mid = (lo + hi) >> 1;
    const uint32_t v = (uint32_t)__shfl((int)rs.incl, (int)mid, 64);
    if (v <= i) lo = mid + 1;
    else hi = mid;
  }
  const uint32_t before = (uint32_t)__shfl((int)(rs.incl - rs.count), (int)lo, 64);
  *pos = i - before;
  return lo;
}

// 6 waves per SIMD for 8-wide rows (80 registers, a few spilled: 3.49 ->
// 3.25 ms per launch of 2^20 traversals at 6-max 100bb against the
// compiler's 124 registers and 4 waves); 16-wide rows keep 4 waves.
template <int ROW, bool TRACE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ROW <= 8 ? 6 : 4))) void k_es_gen(LaunchCfg lc, const DevSeg* __restrict__ segs,
                                                const DevGame* __restrict__ tenants, DevScratch sc, uint32_t g) {
  constexpr int WG = 256;
  __shared__ uint32_t s_sb[kMaxPlayers * WG];
  __shared__ uint32_t s_bk[kMaxPlayers * WG];
  __shared__ uint32_t s_menu[MCCFR_MAX_ACTIONS];
  __shared__ uint32_t s_rcp[MCCFR_MAX_ACTIONS];
  const EsGenBufs& eb = sc.gen;
  const DevSeg& Sg = segs[0];
  const DevGame& G = tenants[uniform(Sg.tenant)];
  // (task counts are final: the previous generation wrote them)
  EsRegionScan ts{0u, 0u, 0u};
  if (g > 0) ts = es_region_scan(eb.task_cnt, g, eb.cap_task_r);
  const uint32_t n_tasks = g == 0 ? Sg.n : ts.total;
  if (n_tasks == 0 || blockIdx.x * WG >= n_tasks) return;  // workgroup-uniform
  const int tx = threadIdx.x, lane = tx & 63;
  if (tx < MCCFR_MAX_ACTIONS) {
    s_menu[tx] = menu_word(G, tx);
    s_rcp[tx] = G.kind[tx] == MCCFR_ACT_POT_FRAC ? den_reciprocal(G.den[tx]) : 0u;
  }
  __syncthreads();  // the last barrier: the returns below need not be uniform
  // an earlier generation overflowed a buffer: the launch is void
  if (uniform(sc.cnt->es_overflow) != 0u) return;
  Rules rules;
  rules.check_mask = G.check_mask;
  rules.call_mask = G.call_mask;
  rules.fold_mask = G.fold_mask;
  rules.allin_mask = G.allin_mask;
  rules.raise_mask = G.raise_mask;
  rules.max_raises = G.max_raises;
  rules.bb = G.bb;
  {
    uint32_t rm = rules.raise_mask;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      rules.fast[i] = 0u;
      if (rm) {
        const int a = __builtin_ctz(rm);
        rm &= rm - 1;
        const bool min_raise = G.kind[a] == MCCFR_ACT_RAISE_MIN;
        rules.fast[i] = (min_raise ? 0u : (uint32_t)G.num[a]) | ((min_raise ? 1u : (uint32_t)G.den[a]) << 8) |
                        ((1u << a) << 16);
      }
    }
    rules.raise_rest = rm;
  }
  const int N = G.N;
  const int hist_bits = G.hist_bits;
  const uint32_t slot_mask = (uint32_t)((1ull << G.table_log2) - 1ull);
  const int up = Sg.update_player;
  const uint32_t allmask = (1u << N) - 1u;
  const size_t L = sc.cap_lanes;
  // the wave's region for the nodes and tasks it makes, and its first free node there
  const uint32_t wreg = uniform(((blockIdx.x * WG + (uint32_t)tx) >> 6) % kEsRegions);
  uint32_t nb = 0;
  for (uint32_t k = 0; k < g; ++k) nb += es_ctr(eb.node_cnt, k, wreg);
  nb = uniform(nb);
  const uint32_t* __restrict__ tin = eb.task[g & 1u];
  uint32_t* __restrict__ tout = eb.task[(g + 1u) & 1u];
  uint32_t* sb = s_sb + tx;
  uint32_t* bk = s_bk + tx;
  const uint64_t lanes_below = (1ull << lane) - 1ull;
  const uint32_t stride = gridDim.x * WG;
  for (uint32_t c0 = uniform(blockIdx.x * WG + (uint32_t)(tx & ~63)); c0 < n_tasks; c0 += stride) {
    const uint32_t i = c0 + (uint32_t)lane;
    const bool live = i < n_tasks;
    uint32_t tref = 0;  // a task of generation >= 1: node << 4 | action, from its region's list
    if (g > 0) {  // (every lane takes part in the region search's shuffles)
      uint32_t pos;
      const uint32_t r = es_region_of(ts, live ? i : 0u, &pos);
      if (live) tref = tin[(size_t)r * eb.cap_task_r + pos];
    }
    uint32_t t = 0, link = kEsNoParent, path = 0, folded = 0, all_in = 0, ab = 0, qmask = 0;
    int qpos = 0, pot = 0, current_bet = 0, lfr = rules.bb, street = 0, raises = 0, depth = 0, dealer = 0;
    uint32_t phash = 0;  // trace: sum of the entered (node, action) pairs' FNV-1a
    DrawStream rng;
    if (live) {
      if (g == 0) {
        // the traversal's root (cfr_gpu.cu:954-981, or the tenant's postflop root), as k_walk
        t = Sg.lane0 + i;
        dealer = (int)sc.root_dealer[t];
        const int root_street = G.root_street;
#pragma unroll
        for (int p = 0; p < kMaxPlayers; ++p) {
          if (p < N) {
            sb[p * WG] = (uint32_t)G.root_stack[p] | ((uint32_t)root_street << 30);
            bk[p * WG] = sc.root_bkt[(size_t)p * L + t];
          }
        }
        if (root_street == 0) {
          const int sb_pos = (N == 2) ? dealer : wrap(dealer + 1, N);
          const int bb_pos = (N == 2) ? wrap(dealer + 1, N) : wrap(dealer + 2, N);
          const int s_sb = (int)(sb[sb_pos * WG] & kStackMask);
          const int sb_amt = G.sb < s_sb ? G.sb : s_sb;
          sb[sb_pos * WG] = (uint32_t)(s_sb - sb_amt) | ((uint32_t)sb_amt << 15);
          if (s_sb - sb_amt == 0) all_in |= 1u << sb_pos;
          const int s_bb = (int)(sb[bb_pos * WG] & kStackMask);
          const int bb_amt = rules.bb < s_bb ? rules.bb : s_bb;
          sb[bb_pos * WG] = (uint32_t)(s_bb - bb_amt) | ((uint32_t)bb_amt << 15);
          if (s_bb - bb_amt == 0) all_in |= 1u << bb_pos;
          pot = sb_amt + bb_amt;
          current_bet = bb_amt;
          qmask = allmask & ~all_in;
          qpos = wrap(bb_pos + 1, N);
        } else {
          folded = G.root_folded;
          all_in = G.root_allin;
          pot = G.root_pot;
          street = root_street;
          ab = G.root_ab;
          qmask = allmask & ~(folded | all_in);
          qpos = wrap(dealer + 1, N);
        }
        rng.init(G.seed_lo, G.seed_hi, Sg.first_tid + i, G.tag);
        rng.k = G.walk_k0;  // the deal (and the preflop dealer draw) came first
      } else {
        // the task's node state, then its action (apply_action_gpu +
        // rebuild_queue, cfr_gpu.cu:543-585, for the update player)
        link = tref;
        const uint4* q = eb.node + (size_t)(link >> 4) * 4;
        const uint4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
        const uint32_t w[kMaxPlayers] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
        t = q2.w;
#pragma unroll
        for (int p = 0; p < kMaxPlayers; ++p) {
          if (p < N) {
            sb[p * WG] = w[p];
            bk[p * WG] = sc.root_bkt[(size_t)p * L + t];  // the turn byte is fetched again if needed
          }
        }
        folded = q1.z & 0xFFu;
        all_in = (q1.z >> 8) & 0xFFu;
        qmask = (q1.z >> 16) & 0xFFu;
        ab = q1.w;
        pot = (int)q2.x;
        current_bet = (int)(q2.y & 0xFFFFu);
        lfr = (int)(q2.y >> 16);
        street = (int)(q2.z & 0xFFu);
        raises = (int)((q2.z >> 8) & 0xFFu);
        depth = (int)((q2.z >> 16) & 0xFFu);
        dealer = (int)(q2.z >> 24);
        const int a = (int)(link & 15u);
        path = es_child_path(q3.y, a);
        if (TRACE) phash += fnv_path(2166136261u, q3.z, (uint32_t)a);
        const uint32_t sbv = sb[up * WG];
        const int stk = (int)(sbv & kStackMask);
        const int bet = ((int)(sbv >> 30) == street) ? (int)((sbv >> 15) & kStackMask) : 0;
        const int to_call = current_bet - bet;
        qmask &= ~(1u << up);
        qpos = wrap(up + 1, N);
        const int chips = chips_lds(G, s_menu, s_rcp, a, pot, stk, to_call, current_bet, lfr);
        ab = ((ab << hist_bits) | (uint32_t)a) & kHistoryMask;
        if (rules.fold_mask >> a & 1u) {
          folded |= 1u << up;
        } else {
          const int ns = stk - chips, nbt = bet + chips;
          sb[up * WG] = (uint32_t)ns | ((uint32_t)nbt << 15) | ((uint32_t)street << 30);
          pot += chips;
          if (ns == 0) all_in |= 1u << up;
          if (nbt > current_bet) {
            const int raise = nbt - current_bet;
            if (raise >= lfr || current_bet == 0) lfr = raise;
            current_bet = nbt;
            ++raises;
            qmask = allmask & ~(folded | all_in) & ~(1u << up);
          }
        }
        ++depth;
        rng.init(G.seed_lo, G.seed_hi, Sg.first_tid + (t - Sg.lane0), G.tag);
        rng.path = path;  // the task's own stream, from draw 0
      }
    }
    // walk to a leaf (status 1 fold, 2 showdown, 3 depth guard) or an update
    // node (status 4); one recursion level of traverse_gpu per pass
    int status = live ? 0 : 9;
    float u = 0.0f;
    uint32_t n_slot = 0, n_vm = 0, dec = 0;
    while (__ballot(status == 0)) {
      if (status != 0) continue;
      if (depth >= kMaxDepth) {  // depth guard (cfr_gpu.cu:661)
        status = 3;
        continue;
      }
      const uint32_t alive = allmask & ~folded;
      if (__popc(alive) == 1) {
        const int inv_up = G.stack[up] - (int)(sb[up * WG] & kStackMask);
        u = ((int)__builtin_ctz(alive) == up) ? (float)(pot - inv_up) : (float)(-inv_up);
        status = 1;
        continue;
      }
      if (qmask == 0) {
        // street change (cfr_gpu.cu:668-692), then the next level's decision in
        // this pass; with nobody left to act, or after the river, the showdown
        if (street == 3) {
          status = 2;
          continue;
        }
        ++street;
        current_bet = 0;
        lfr = rules.bb;
        raises = 0;
        qmask = allmask & ~(folded | all_in);
        qpos = wrap(dealer + 1, N);
        if (qmask == 0) {
          status = 2;
          continue;
        }
        ++depth;
        if (depth >= kMaxDepth) {
          status = 3;
          continue;
        }
      }
      const uint32_t rot = ((qmask >> qpos) | (qmask << (N - qpos))) & allmask;
      const int player = wrap(qpos + (int)__builtin_ctz(rot), N);
      const uint32_t sbv = sb[player * WG];
      const int stk = (int)(sbv & kStackMask);
      const int bet = ((int)(sbv >> 30) == street) ? (int)((sbv >> 15) & kStackMask) : 0;
      const int to_call = current_bet - bet;
      const uint32_t vm = rules_mask(rules, s_menu, pot, stk, to_call, current_bet, lfr, raises);
      uint32_t Bp = bk[player * WG];
      if (street == 2 && ((Bp >> 16) & 0xFFu) == kTurnUnknown) {  // first turn decision of this seat
        const uint32_t tb = sc.turn6[sc.root_tidx[(size_t)player * L + t]];
        Bp = (Bp & 0xFF00FFFFu) | (tb << 16);
        bk[player * WG] = Bp;
      }
      const uint32_t bbk = street == 0 ? 0u : (Bp >> (8 * street)) & 0xFFu;
      const uint32_t slot = info_hash((uint32_t)player, Bp & 0xFFu, bbk, (uint32_t)street, ab, slot_mask);
      ++dec;
      if (player == up) {
        if (g + 1u >= kEsGens) {
          // cannot happen (see kEsGens); flagged rather than left unwalked
          sc.cnt->es_overflow = 2u;
          status = 3;
          continue;
        }
        status = 4;
        n_slot = slot;
        n_vm = vm;
        continue;
      }
      // sample_action (cfr_gpu.cu:510-537) on the BF16 snapshot, as k_walk
      float sg[ROW];
      load_bf16_row<ROW>(G.sig + (size_t)slot * ROW, sg);
      float cum[ROW];
      float total = 0.0f;
#pragma unroll
      for (int a = 0; a < ROW; ++a) {
        total += (vm >> a & 1u) ? sg[a] : 0.0f;
        cum[a] = total;
      }
      int act = -1;
      if (total >= 1e-7f) {
        const float r = u01(rng.next()) * total;
        int below = 0;
#pragma unroll
        for (int a = 0; a < ROW; ++a) below += cum[a] < r ? 1 : 0;
        act = below;
      }
      if (act < 0) {  // passive fallback (cfr_gpu.cu:498-508)
        for (int k = 0; k < G.n_passive; ++k) {
          const int a = (int)(G.passive_order[k] & 15u);
          if (vm >> a & 1u) {
            act = a;
            break;
          }
        }
        if (act < 0) act = 0;
      }
      if (TRACE) phash += fnv_path(2166136261u, slot, (uint32_t)act);
      // apply_action_gpu + rebuild_queue (cfr_gpu.cu:543-585)
      qmask &= ~(1u << player);
      qpos = wrap(player + 1, N);
      const int chips = chips_lds(G, s_menu, s_rcp, act, pot, stk, to_call, current_bet, lfr);
      ab = ((ab << hist_bits) | (uint32_t)act) & kHistoryMask;
      if (rules.fold_mask >> act & 1u) {
        folded |= 1u << player;
      } else {
        const int ns = stk - chips, nbt = bet + chips;
        sb[player * WG] = (uint32_t)ns | ((uint32_t)nbt << 15) | ((uint32_t)street << 30);
        pot += chips;
        if (ns == 0) all_in |= 1u << player;
        if (nbt > current_bet) {
          const int raise = nbt - current_bet;
          if (raise >= lfr || current_bet == 0) lfr = raise;
          current_bet = nbt;
          ++raises;
          qmask = allmask & ~(folded | all_in) & ~(1u << player);
        }
      }
      ++depth;
    }
    const bool leaf = status >= 1 && status <= 3;
    if (status == 2) {  // showdown payoff (cfr_gpu.cu:694-752)
      int inv[kMaxPlayers];
      uint32_t rkv[kMaxPlayers];
#pragma unroll
      for (int p = 0; p < kMaxPlayers; ++p) {
        inv[p] = p < N ? G.stack[p] - (int)(sb[p * WG] & kStackMask) : 0;
        rkv[p] = p < N ? (uint32_t)sc.root_rank[(size_t)p * L + t] : 0u;
      }
      u = showdown_payoff(up, folded, inv, rkv);
    }
    if (leaf) {  // status 3: the depth guard's 0 (cfr_gpu.cu:661)
      if (link != kEsNoParent)
        sc.es_rd[(size_t)(link >> 4) * ROW + (link & 15u)] = u;
      else if (TRACE)
        eb.root_v[t] = u;
    }
    // update nodes: an index each, then one task per legal action
    const bool isnode = status == 4;
    const uint64_t nmask = __ballot(isnode);
    // (indices within the wave's region wreg; node = wreg * cap_node_r + local)
    uint32_t node = 0xFFFFFFFFu, nch = 0, cfirst = 0;
    bool stored = false;
    if (nmask) {
      uint32_t nbase = 0;
      if (lane == 0) nbase = atomicAdd(&es_ctr(eb.node_cnt, g, wreg), (uint32_t)__popcll(nmask));
      nbase = __shfl(nbase, 0, 64);
      if (isnode) {
        const uint32_t loc = nb + nbase + (uint32_t)__popcll(nmask & lanes_below);
        stored = loc < eb.cap_node_r;
        node = wreg * eb.cap_node_r + loc;
        nch = stored ? (uint32_t)__popc(n_vm) : 0u;
        if (!stored) sc.cnt->es_overflow = 1u;
      }
      uint32_t incl = nch;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
      }
      const uint32_t total = __shfl(incl, 63, 64);
      uint32_t cbase = 0;
      if (lane == 0 && total) cbase = atomicAdd(&es_ctr(eb.task_cnt, g + 1, wreg), total);
      cbase = __shfl(cbase, 0, 64);
      cfirst = wreg * eb.cap_task_r + cbase + incl - nch;
      if (total && cbase + total > eb.cap_task_r) {
        if (lane == 0) sc.cnt->es_overflow = 3u;
        nch = 0;  // the next generation does not run
      }
    }
    if (stored) {
      // the node's state before its action, then a task per legal action
      uint32_t w[kMaxPlayers];
#pragma unroll
      for (int p = 0; p < kMaxPlayers; ++p) w[p] = p < N ? sb[p * WG] : 0u;
      uint4* q = eb.node + (size_t)node * 4;
      q[0] = make_uint4(w[0], w[1], w[2], w[3]);
      q[1] = make_uint4(w[4], w[5], folded | (all_in << 8) | (qmask << 16) | ((uint32_t)qpos << 24), ab);
      q[2] = make_uint4((uint32_t)pot, (uint32_t)current_bet | ((uint32_t)lfr << 16),
                        (uint32_t)street | ((uint32_t)raises << 8) | ((uint32_t)depth << 16) | ((uint32_t)dealer << 24),
                        t);
      q[3] = make_uint4(link, path, n_slot, n_vm);
      uint32_t ci = cfirst;
      for (uint32_t left = nch ? n_vm : 0u; left; left &= left - 1u) tout[ci++] = node << 4 | (uint32_t)__builtin_ctz(left);
    }
    if (TRACE && live) {
      // per-traversal sums for the trace: the root task writes them first
      uint32_t* acc = eb.tr_acc + (size_t)t * 8;
      const uint32_t upd = isnode ? 1u : 0u, ul = isnode ? (uint32_t)__popc(n_vm) : 0u, lv = leaf ? 1u : 0u;
      const uint32_t tb = leaf ? 1u << status : 0u;
      if (g == 0) {
        acc[0] = dec;
        acc[1] = upd;
        acc[2] = ul;
        acc[3] = lv;
        acc[4] = tb;
        acc[5] = phash;
      } else {
        atomicAdd(acc + 0, dec);
        atomicAdd(acc + 1, upd);
        atomicAdd(acc + 2, ul);
        atomicAdd(acc + 3, lv);
        atomicOr(acc + 4, tb);
        atomicAdd(acc + 5, phash);
      }
    }
  }
}

// Generation g's update nodes, one lane each: v(I) = sum_a sigma(a) v(a)
// (sigma from the FP32 regrets, in action order, as k_walk_es), the record
// (at the node's index) and its row of v(a) - v(I), and v(I) into the
// parent's child slot (or the traversal's root value).
template <int ROW, bool TRACE>
__global__ __launch_bounds__(256) void k_es_back(LaunchCfg lc, const DevSeg* __restrict__ segs,
                                                 const DevGame* __restrict__ tenants, DevScratch sc, uint32_t g) {
  const EsGenBufs& eb = sc.gen;
  if (uniform(sc.cnt->es_overflow) != 0u) return;  // (no barrier in this kernel)
  const EsRegionScan ns = es_region_scan(eb.node_cnt, g, eb.cap_node_r);
  const uint32_t n = ns.total;
  if (n == 0 || blockIdx.x * blockDim.x >= n) return;
  // lane r: region r's nodes of earlier generations (where generation g's
  // start there); their sum is generation g's first record
  const int lane = threadIdx.x & 63;
  uint32_t before = 0;
  for (uint32_t k = 0; k < g; ++k) before += es_ctr(eb.node_cnt, k, (uint32_t)lane);
  uint32_t rec0 = before;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) rec0 += __shfl_xor(rec0, d, 64);
  rec0 = uniform(rec0);
  const DevGame& G = tenants[uniform(segs[0].tenant)];
  const int A = G.num_actions;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t c0 = uniform(blockIdx.x * blockDim.x + (threadIdx.x & ~63u)); c0 < n; c0 += stride) {
    const uint32_t j = c0 + (uint32_t)lane;
    const bool live = j < n;
    uint32_t pos;
    const uint32_t r = es_region_of(ns, live ? j : 0u, &pos);
    const uint32_t first = (uint32_t)__shfl((int)before, (int)r, 64);
    if (!live) continue;
    const uint32_t node = r * eb.cap_node_r + first + pos;  // below the region's capacity: no overflow
    const uint4 q3 = eb.node[(size_t)node * 4 + 3];
    const uint32_t slot = q3.z, vm = q3.w;
    float rr[ROW], v[ROW], rd[ROW];
    load_f32_row<ROW>(G.R + (size_t)slot * 2 * ROW, rr);
    load_f32_row<ROW>(sc.es_rd + (size_t)node * ROW, v);  // child values (legal actions)
    float sum_r = 0.0f;
#pragma unroll
    for (int a = 0; a < ROW; ++a) {
      const bool legal = a < A && (vm >> a & 1u);
      rr[a] = legal ? fmaxf(0.0f, rr[a]) : 0.0f;
      sum_r += rr[a];
      v[a] = legal ? v[a] : 0.0f;
    }
    const float uni = 1.0f / (float)__popc(vm);
    float vi = 0.0f;
#pragma unroll
    for (int a = 0; a < ROW; ++a) {
      const bool legal = a < A && (vm >> a & 1u);
      const float sig = sum_r > 1e-7f ? rr[a] / sum_r : (legal ? uni : 0.0f);
      if (legal) vi += sig * v[a];
    }
#pragma unroll
    for (int a = 0; a < ROW; ++a) rd[a] = (vm >> a & 1u) ? v[a] - vi : 0.0f;
    store_f32_row<ROW>(sc.es_rd + (size_t)node * ROW, rd);
    // the node's record, at generation g's records' next place (one segment: key = slot)
    const uint32_t rec = rec0 + j;
    sc.key[0][rec] = slot;
    sc.val[0][rec] = make_uint2((vm << 8) | kEsRecord, node);
    const uint32_t link = q3.x;
    if (link != kEsNoParent)
      sc.es_rd[(size_t)(link >> 4) * ROW + (link & 15u)] = vi;
    else if (TRACE)
      eb.root_v[eb.node[(size_t)node * 4 + 2].w] = vi;
  }
}

// The launch's record count (one per node) for the sort.
__global__ __launch_bounds__(64) void k_es_gen_end(DevScratch sc) {
  const uint32_t r = threadIdx.x;  // one region per lane
  uint32_t c = 0;
  for (uint32_t k = 0; k <= kEsGens; ++k) c += min(es_ctr(sc.gen.node_cnt, k, r), sc.gen.cap_node_r);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
  if (r == 0) {
    sc.cnt->n_es = c;
    sc.cnt->n_es_valid = c;
  }
}

// Trace records of a one-segment generation launch.
__global__ __launch_bounds__(256) void k_es_gen_trace(const DevSeg* __restrict__ segs, DevScratch sc,
                                                      mccfr_trace* __restrict__ trace) {
  const DevSeg& Sg = segs[0];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Sg.n) return;
  const uint32_t t = Sg.lane0 + i;
  const uint32_t* acc = sc.gen.tr_acc + (size_t)t * 8;
  mccfr_trace r;
  r.utility = sc.gen.root_v[t];
  r.path_hash = 2166136261u + acc[5];
  r.decisions = (uint16_t)acc[0];
  r.updates = (uint8_t)acc[1];
  r.terminal = (uint8_t)acc[4];
  r.update_legal = (uint16_t)acc[2];
  r.reserved = (uint16_t)acc[3];
  trace[t] = r;
}

}  // namespace

void launch_walk(hipStream_t s, int row, const LaunchCfg& lc, const DevSeg* segs, const DevGame* tenants,
                 const DealInputs& in, const DevScratch& sc, WalkShape shape, mccfr_trace* trace,
                 hipEvent_t after_deal, hipEvent_t after_walk) {
  hipLaunchKernelGGL(k_deal, dim3(lc.lanes / kDealWG), dim3(kDealWG), 0, s, lc, segs, tenants, in, sc);
  if (after_deal) (void)hipEventRecord(after_deal, s);
  DevScratch ws = sc;
  ws.turn6 = in.bt.turn6;
  if (lc.es && lc.n_segs == 1 && sc.gen.task[0]) {
    // external sampling by generations: each generation's tasks walk to
    // their first update node or leaf; then the nodes' values, deepest
    // generation first
    hipLaunchKernelGGL(k_es_gen_begin, dim3(1), dim3(256), 0, s, ws);
    const uint32_t g0 = (lc.lanes + 255u) / 256u;
    const uint32_t gn = std::max<uint32_t>(1u, std::min<uint32_t>(4096u, (lc.lanes * 2u + 255u) / 256u));
    for (uint32_t g = 0; g < kEsGens; ++g) {
      const dim3 grid(g == 0 ? g0 : gn);
#define MCCFR_GEN(R, T) hipLaunchKernelGGL((k_es_gen<R, T>), grid, dim3(256), 0, s, lc, segs, tenants, ws, g)
      if (row == 8) {
        if (trace) MCCFR_GEN(8, true);
        else MCCFR_GEN(8, false);
      } else {
        if (trace) MCCFR_GEN(16, true);
        else MCCFR_GEN(16, false);
      }
#undef MCCFR_GEN
    }
    for (uint32_t g = kEsGens; g-- > 0;) {
#define MCCFR_BACK(R, T) hipLaunchKernelGGL((k_es_back<R, T>), dim3(gn), dim3(256), 0, s, lc, segs, tenants, ws, g)
      if (row == 8) {
        if (trace) MCCFR_BACK(8, true);
        else MCCFR_BACK(8, false);
      } else {
        if (trace) MCCFR_BACK(16, true);
        else MCCFR_BACK(16, false);
      }
#undef MCCFR_BACK
    }
    hipLaunchKernelGGL(k_es_gen_end, dim3(1), dim3(64), 0, s, ws);
    if (trace) hipLaunchKernelGGL(k_es_gen_trace, dim3(g0), dim3(256), 0, s, segs, ws, trace);
    if (after_walk) (void)hipEventRecord(after_walk, s);
    return;
  }
  if (lc.es) {
    // external sampling over several segments: one lane per traversal, depth
    // first, showdowns settled in the walk
    hipLaunchKernelGGL(k_es_begin, dim3(1), dim3(64), 0, s, sc);
    const dim3 grid(es_grid(lc.lanes));
#define MCCFR_ES(R, T) \
  hipLaunchKernelGGL((k_walk_es<R, T>), grid, dim3(256), 0, s, lc, segs, tenants, ws, sc.es_stack, trace)
    if (row == 8) {
      if (trace) MCCFR_ES(8, true);
      else MCCFR_ES(8, false);
    } else {
      if (trace) MCCFR_ES(16, true);
      else MCCFR_ES(16, false);
    }
#undef MCCFR_ES
    if (after_walk) (void)hipEventRecord(after_walk, s);
    return;
  }
  if (row == 8)
    walk_dispatch<8>(s, lc, segs, tenants, ws, shape, trace);
  else
    walk_dispatch<16>(s, lc, segs, tenants, ws, shape, trace);
  if (after_walk) (void)hipEventRecord(after_walk, s);
  // the records, in lane order, with the showdowns' payoffs
  hipLaunchKernelGGL(k_emit_scan, dim3(1), dim3(1024), 0, s, lc, segs, tenants, sc);
  const dim3 grid((lc.lanes + kEmitLanes - 1) / kEmitLanes);
  if (trace)
    hipLaunchKernelGGL(k_emit<true>, grid, dim3(256), 0, s, lc, segs, tenants, sc, trace);
  else
    hipLaunchKernelGGL(k_emit<false>, grid, dim3(256), 0, s, lc, segs, tenants, sc, trace);
}

size_t es_stack_bytes(int row, uint32_t lanes) {
  const int words = row == 8 ? es_frame_words<8>() : es_frame_words<16>();
  return (size_t)kEsFrames * (size_t)words * es_grid(lanes) * 256u * sizeof(uint32_t);
}

size_t es_gen_bytes(uint32_t cap_lanes) {
  const auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t cap_r = (size_t)cap_lanes * kMaxDepth / kEsRegions;  // nodes and tasks per region
  const size_t ctrs = (size_t)(kEsGens + 1) * kEsRegions * kEsCtrStride;
  return 2 * al(kEsRegions * cap_r * sizeof(uint32_t)) + 2 * al(ctrs * sizeof(uint32_t)) +
         al(kEsRegions * cap_r * sizeof(uint4) * 4) + al((size_t)cap_lanes * sizeof(float)) +
         al((size_t)cap_lanes * 8 * sizeof(uint32_t));
}

void carve_es_gen(EsGenBufs& b, char* base, uint32_t cap_lanes) {
  const auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t cap_r = (size_t)cap_lanes * kMaxDepth / kEsRegions;  // nodes and tasks per region
  const size_t ctrs = (size_t)(kEsGens + 1) * kEsRegions * kEsCtrStride;
  char* p = base;
  const auto take = [&](size_t bytes) {
    char* q = p;
    p += al(bytes);
    return q;
  };
  b.task[0] = reinterpret_cast<uint32_t*>(take(kEsRegions * cap_r * sizeof(uint32_t)));
  b.task[1] = reinterpret_cast<uint32_t*>(take(kEsRegions * cap_r * sizeof(uint32_t)));
  b.task_cnt = reinterpret_cast<uint32_t*>(take(ctrs * sizeof(uint32_t)));
  b.node_cnt = reinterpret_cast<uint32_t*>(take(ctrs * sizeof(uint32_t)));
  b.node = reinterpret_cast<uint4*>(take(kEsRegions * cap_r * sizeof(uint4) * 4));
  b.root_v = reinterpret_cast<float*>(take((size_t)cap_lanes * sizeof(float)));
  b.tr_acc = reinterpret_cast<uint32_t*>(take((size_t)cap_lanes * 8 * sizeof(uint32_t)));
  b.cap_task_r = (uint32_t)cap_r;
  b.cap_node_r = (uint32_t)cap_r;
}


void launch_reduce(hipStream_t s, int row, const LaunchCfg& lc, const DevSeg* segs, const DevGame* tenants,
                   const DevScratch& sc, uint32_t sort_bits, DevSeg* advance) {
  // (outcome-sampling records were emitted after the walk; external-sampling
  // walks append theirs, and the scan only sets up the update)
  if (lc.es) hipLaunchKernelGGL(k_emit_scan, dim3(1), dim3(1024), 0, s, lc, segs, tenants, sc);
  // at most lanes * kMaxDepth records (external sampling: up to the sort
  // input's capacity); the sort kernels loop over the tiles
  const uint64_t max_rec = (uint64_t)(lc.es ? sc.cap_lanes : lc.lanes) * kMaxDepth;
  const uint32_t max_rec_tiles = (uint32_t)((max_rec + kSortTile - 1) / kSortTile);
  // a multiple of the XCD count (the tile order assumes it)
  const uint32_t sgrid = std::min<uint32_t>((std::max<uint32_t>(max_rec_tiles, 1u) + 7u) / 8u * 8u, 2048u);
  // as few passes of at most kDigitBits as the sorted bits need, spread
  // evenly over them (26 bits: 3 passes of 9 bits, 512 digits each)
  const uint32_t passes = sort_bits == 0 ? 1u : (sort_bits + kDigitBits - 1u) / kDigitBits;
  const uint32_t bits = std::max(kMinDigitBits, std::min(kDigitBits, (sort_bits + passes - 1u) / passes));
  for (uint32_t p = 0; p < passes; ++p) {
    const uint32_t shift = bits * p;
    const uint32_t* kin = sc.key[p & 1];
    const uint2* vin = sc.val[p & 1];
#define MCCFR_PASS(B)                                                                                     \
  hipLaunchKernelGGL(k_rs_hist<B>, dim3(sgrid), dim3(kSortWG), 0, s, kin, sc.cnt, shift, sc.hist, sc.max_tiles);  \
  hipLaunchKernelGGL(k_rs_scan, dim3(1u << B), dim3(256), 0, s, sc.hist, sc.bin_total, sc.cnt, sc.max_tiles);   \
  hipLaunchKernelGGL(k_rs_scatter<B>, dim3(sgrid), dim3(kSortWG), 0, s, kin, vin, sc.key[(p + 1) & 1],            \
                     sc.val[(p + 1) & 1], sc.cnt, shift, sc.hist, sc.bin_total, sc.max_tiles)
    if (bits == 8) {
      MCCFR_PASS(8);
    } else if (bits == 9) {
      MCCFR_PASS(9);
    } else {
      MCCFR_PASS(10);
    }
#undef MCCFR_PASS
  }
  const uint32_t* keys = sc.key[passes & 1];
  const uint2* vals = sc.val[passes & 1];
  // (at most 2,048 workgroups: a wave's 64 records are little work, and a grid
  // sized for the record bound dispatches mostly empty waves; 8,192 took
  // 9.0 us per launch at 6-max 100bb against 7.6)
  const uint32_t rgrid = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((max_rec / 4 + 255) / 256, 1), 2048);
  hipLaunchKernelGGL(k_runs, dim3(rgrid), dim3(256), 0, s, sc, keys, advance, tenants, lc.n_segs);
  // the long-run search's workgroups lead the short runs' grid.  The reduce
  // grids (4 waves per workgroup, all grid-stride) are sized from the record
  // bound, so a small launch does not dispatch thousands of empty workgroups
  const uint32_t cand_blocks = rgrid / 64 + 1;
  const auto wg4 = [](uint64_t items, uint32_t cap) {
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>((items + 3) / 4, 1), cap);
  };
  const uint64_t max_long = max_rec / (kShortRun + 1) + 1;
  const uint32_t rows_grid = wg4((max_rec + 127) / 128, 8192);  // a wave per pair of 64-record regions
  const uint32_t piece_grid = wg4(max_rec / kPiece + max_long, 8192);
  const uint32_t runs_grid = wg4(max_long, 2048);
  // the record kinds are known per launch (a launch is all external
  // sampling or none), so outcome-sampling launches run reduce kernels
  // without the delta-row path
#define MCCFR_REDUCE(R, E)                                                                                    \
  hipLaunchKernelGGL((k_reduce_rows<R, E>), dim3(cand_blocks + rows_grid), dim3(256), 0, s, lc, sc, keys, vals, \
                     cand_blocks);                                                                            \
  hipLaunchKernelGGL((k_reduce_piece<R, E>), dim3(piece_grid), dim3(256), 0, s, lc, sc, keys, vals);           \
  hipLaunchKernelGGL(k_reduce_runs<R>, dim3(runs_grid), dim3(256), 0, s, lc, sc, keys)
  if (row == 8) {
    if (lc.es) {
      MCCFR_REDUCE(8, true);
    } else {
      MCCFR_REDUCE(8, false);
    }
  } else {
    if (lc.es) {
      MCCFR_REDUCE(16, true);
    } else {
      MCCFR_REDUCE(16, false);
    }
  }
#undef MCCFR_REDUCE
}

}  // namespace mccfr
