// The traversal pipeline of one launch (the reference's simulate kernel,
// cfr_gpu.cu:878-1011, and its clamp / regret-matching sweeps, :1018-1022 and
// :318-351), as four gfx950 kernels:
//
//   k_deal      one lane per traversal: Philox deal, hole and board buckets
//               (precomputed bucket tables), showdown ranks, dealer -> a
//               compact root record in HBM.
//   k_walk      persistent lanes walk traversal after traversal from the root
//               records: both node types follow exactly one child
//               (cfr_gpu.cu:800, :844), so a traversal is one root-to-leaf
//               path; lanes keep their game state in LDS and record only the
//               update-player nodes (slot, sampled action, legal mask) and the
//               terminal utility.
//   k_scatter   every recorded node's regret and strategy-sum contributions
//               (cfr_gpu.cu:784-837) added into FP64 per-launch deltas.
//   k_update    each touched slot once: regrets += deltas (CFR+ clamp),
//               strategy sums += deltas, BF16 snapshot re-matched, deltas
//               cleared.  Untouched slots need no sweep: clamp and regret
//               matching are per-slot and idempotent on them.
#include <hip/hip_runtime.h>

#include "kernel_util.hpp"
#include "kernels.hpp"

namespace mccfr {
namespace {

constexpr int kDealWG = 256;
constexpr int kWalkWG = 256;

// ---------------------------------------------------------------------------
// k_deal.  Buckets come from precomputed tables instead of the reference's
// per-player table walks: the rank-walk word after 5 (6) cards depends only on
// the card set, so the flop (turn) bucket of every 5- (6-) card set is stored
// once, indexed by colex rank (kernels.hip k_build_buckets); the river rank and
// bucket come from the arithmetic evaluator.  Per player: three independent
// loads instead of a chain of dependent gathers into the 130 MB table.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kDealWG) void k_deal(const DevGame* __restrict__ gp, const LaunchState* __restrict__ st,
                                                  BucketTables bt, ArithTables at, const uint8_t* __restrict__ lut_g,
                                                  DevTables T, uint32_t n_lanes) {
  __shared__ uint32_t s_lut[2704 / 4];
  __shared__ uint16_t s_off[13 * 8 * 5];
  __shared__ uint8_t s_deck[kDeckStride * kDealWG];
  const DevGame& G = *gp;
  for (int i = threadIdx.x; i < 2704 / 4; i += kDealWG) s_lut[i] = reinterpret_cast<const uint32_t*>(lut_g)[i];
  for (int i = threadIdx.x; i < 13 * 8 * 5; i += kDealWG) s_off[i] = at.offsets[i];
  __syncthreads();
  uint32_t n = (uint32_t)st->n;
  if (n > n_lanes) n = n_lanes;
  const uint32_t t = blockIdx.x * kDealWG + threadIdx.x;
  if (t >= n) return;
  const int N = G.N;

  // partial Fisher-Yates over 2N+5 draws (cfr_gpu.cu:460-470, 931-937)
  DrawStream rng;
  rng.init(G.seed_lo, G.seed_hi, st->first_tid + t, G.tag);
  uint8_t* dk = s_deck + threadIdx.x * kDeckStride;
  for (int i = 0; i < 52; ++i) dk[i] = (uint8_t)i;
  const int n_deal = 2 * N + 5;
  for (int i = 0; i < n_deal; ++i) {
    const uint32_t r = rng.next();
    const int j = i + (int)(((uint64_t)r * (uint64_t)(52 - i)) >> 32);
    const uint8_t a = dk[i], b = dk[j];
    dk[i] = b;
    dk[j] = a;
  }
  const uint32_t f0 = dk[2 * N], f1 = dk[2 * N + 1], f2 = dk[2 * N + 2], c3 = dk[2 * N + 3], c4 = dk[2 * N + 4];
  const uint8_t* lut = reinterpret_cast<const uint8_t*>(s_lut);
#pragma unroll
  for (int pl = 0; pl < kMaxPlayers; ++pl) {
    if (pl < N) {
      const uint32_t h0 = dk[2 * pl], h1 = dk[2 * pl + 1];
      const uint32_t hb = lut[h0 * 52 + h1];
      const uint32_t bf = bt.flop5[colex5(h0, h1, f0, f1, f2)];
      const uint32_t btn = bt.turn6[colex6(h0, h1, f0, f1, f2, c3)];
      const uint32_t c7[7] = {h0, h1, f0, f1, f2, c3, c4};
      const uint32_t r7 = arith_rank7(c7, at.nonflush, at.flush, s_off);
      T.root_bkt[(size_t)pl * T.lanes + t] = hb | (bf << 8) | (btn << 16) | (bucket_of_rank(r7) << 24);
      T.root_rank[(size_t)pl * T.lanes + t] = (uint16_t)r7;
    }
  }
  T.root_dealer[t] = (uint8_t)(rng.next() % (uint32_t)N);  // draw 2N+5 (cfr_gpu.cu:963)
}

// chips_for over the LDS copy of the menu (divergent action index).
__device__ __forceinline__ int chips_lds(const DevGame& g, const uint8_t* kind, const uint8_t* num,
                                         const uint8_t* den, int a, int pot, int stack, int to_call,
                                         int current_bet, int lfr) {
  if (pot <= 0) pot = 1;
  const int k = kind[a];
  if (k == MCCFR_ACT_FOLD || k == MCCFR_ACT_CHECK) return 0;
  if (k == MCCFR_ACT_CALL || k == MCCFR_ACT_CHECK_CALL) return to_call < stack ? to_call : stack;
  if (k == MCCFR_ACT_ALL_IN) return stack;
  const int mrt = min_raise_to(current_bet, lfr, g.bb);
  const int head = stack - to_call;
  const int maxrt = current_bet + (head > 0 ? head : 0);
  int target = mrt;
  if (k == MCCFR_ACT_POT_FRAC) {
    int inc = ((int)num[a] * pot) / (int)den[a];
    inc = inc > 1 ? inc : 1;
    target = current_bet + inc > mrt ? current_bet + inc : mrt;
  }
  target = target < maxrt ? target : maxrt;
  const int add = target - current_bet;
  const int chips = to_call + (add > 0 ? add : 0);
  return chips < stack ? chips : stack;
}

// ---------------------------------------------------------------------------
// k_walk
// ---------------------------------------------------------------------------
template <int ROW>
__global__ __launch_bounds__(kWalkWG) void k_walk(const DevGame* __restrict__ gp, const LaunchState* __restrict__ st,
                                                  DevTables T, uint32_t n_lanes, mccfr_trace* __restrict__ trace) {
  // Per-lane state, field-major ([field][lane]): any per-lane field index
  // maps consecutive lanes to consecutive banks.
  __shared__ uint32_t s_sb[kMaxPlayers * kWalkWG];    // stack | bet << 16
  __shared__ uint32_t s_bk[kMaxPlayers * kWalkWG];    // buckets: hole | flop << 8 | turn << 16 | river << 24
  __shared__ uint16_t s_rk[kMaxPlayers * kWalkWG];    // 7-card rank at showdown
  __shared__ uint32_t s_ps[kRegPending * kWalkWG];    // pending update nodes: slot
  __shared__ uint32_t s_pa[kRegPending * kWalkWG];    //                       action | legal mask << 8
  __shared__ uint8_t s_kind[MCCFR_MAX_ACTIONS], s_num[MCCFR_MAX_ACTIONS], s_den[MCCFR_MAX_ACTIONS];
  const DevGame& G = *gp;
  if (threadIdx.x < MCCFR_MAX_ACTIONS) {
    s_kind[threadIdx.x] = G.kind[threadIdx.x];
    s_num[threadIdx.x] = G.num[threadIdx.x];
    s_den[threadIdx.x] = G.den[threadIdx.x];
  }
  __syncthreads();

  const int tx = threadIdx.x;
  uint32_t* sb = s_sb + tx;
  uint32_t* bk = s_bk + tx;
  uint16_t* rk = s_rk + tx;
  uint32_t* ps = s_ps + tx;
  uint32_t* pa = s_pa + tx;

  uint32_t n = (uint32_t)st->n;
  if (n > n_lanes) n = n_lanes;
  const uint32_t stride = gridDim.x * kWalkWG;
  const int up = st->update_player;
  const uint64_t tid0 = st->first_tid;
  const int N = G.N;
  const int A = G.num_actions;
  const uint32_t slot_mask = (uint32_t)((1ull << G.table_log2) - 1ull);
  const uint32_t allmask = (1u << N) - 1u;
  const size_t L = T.lanes;

  uint32_t t = blockIdx.x * kWalkWG + tx;
  bool live = t < n;

  DrawStream rng;
  uint32_t folded = 0, all_in = 0, ab = 0, queue = 0, phash = 0;
  int pot = 0, current_bet = 0, lfr = 0, street = 0, raises = 0, depth = 0, nq = 0, dealer = 0;
  int npend = 0, decisions = 0, upd_legal = 0;

  // Root state of traversal t (cfr_gpu.cu:954-981): stacks, blinds, preflop queue.
  auto init = [&]() {
#pragma unroll
    for (int p = 0; p < kMaxPlayers; ++p) {
      if (p < N) {
        sb[p * kWalkWG] = (uint32_t)G.stack[p];
        bk[p * kWalkWG] = T.root_bkt[(size_t)p * L + t];
        rk[p * kWalkWG] = T.root_rank[(size_t)p * L + t];
      }
    }
    dealer = T.root_dealer[t];
    const int sb_pos = (N == 2) ? dealer : wrap(dealer + 1, N);
    const int bb_pos = (N == 2) ? wrap(dealer + 1, N) : wrap(dealer + 2, N);
    folded = 0;
    all_in = 0;
    const int s_sb = (int)(sb[sb_pos * kWalkWG] & 0xFFFFu);
    const int sb_amt = G.sb < s_sb ? G.sb : s_sb;
    sb[sb_pos * kWalkWG] = (uint32_t)(s_sb - sb_amt) | ((uint32_t)sb_amt << 16);
    if (s_sb - sb_amt == 0) all_in |= 1u << sb_pos;
    const int s_bb = (int)(sb[bb_pos * kWalkWG] & 0xFFFFu);
    const int bb_amt = G.bb < s_bb ? G.bb : s_bb;
    sb[bb_pos * kWalkWG] = (uint32_t)(s_bb - bb_amt) | ((uint32_t)bb_amt << 16);
    if (s_bb - bb_amt == 0) all_in |= 1u << bb_pos;
    pot = sb_amt + bb_amt;
    current_bet = bb_amt;
    lfr = G.bb;
    street = 0;
    raises = 0;
    ab = 0;
    depth = 0;
    queue = 0;
    nq = 0;
    for (int i = 0, p = wrap(bb_pos + 1, N); i < N; ++i, p = wrap(p + 1, N))
      if (!(all_in >> p & 1u)) { queue |= (uint32_t)p << (3 * nq); ++nq; }
    rng.init(G.seed_lo, G.seed_hi, tid0 + t, G.tag);
    rng.k = (uint32_t)(2 * N + 6);  // the deal and the dealer draw came first
    npend = 0;
    decisions = 0;
    upd_legal = 0;
    phash = 2166136261u;
  };

  auto finish = [&](float u, int term) {
    if (trace) {
      mccfr_trace r;
      r.utility = u;
      r.path_hash = phash;
      r.decisions = (uint16_t)decisions;
      r.updates = (uint8_t)npend;
      r.terminal = (uint8_t)term;
      r.update_legal = (uint16_t)upd_legal;
      r.reserved = 0;
      trace[t] = r;
    }
    T.node_count[t] = (uint8_t)npend;
    T.node_u[t] = u;
    const int kr = npend < kRegPending ? npend : kRegPending;
    for (int k = 0; k < kr; ++k) {
      T.node_slot[(size_t)k * L + t] = ps[k * kWalkWG];
      T.node_av[(size_t)k * L + t] = pa[k * kWalkWG];
    }
    t += stride;
    live = t < n;
    if (live) init();
  };

  if (live) init();
  while (!__all(!live)) {
    if (!live) continue;
    // one recursion level of traverse_gpu (cfr_gpu.cu:644-860)
    if (depth >= kMaxDepth) {
      finish(0.0f, 3);
      continue;
    }
    const uint32_t alive = allmask & ~folded;
    if (__popc(alive) == 1) {
      const int inv_up = G.stack[up] - (int)(sb[up * kWalkWG] & 0xFFFFu);
      finish(((int)__builtin_ctz(alive) == up) ? (float)(pot - inv_up) : (float)(-inv_up), 1);
      continue;
    }
    if (nq == 0) {
      if (street < 3) {
        ++street;
        for (int p = 0; p < N; ++p) sb[p * kWalkWG] &= 0xFFFFu;
        current_bet = 0;
        lfr = G.bb;
        raises = 0;
        for (int i = 0, p = wrap(dealer + 1, N); i < N; ++i, p = wrap(p + 1, N))
          if (!((folded | all_in) >> p & 1u)) { queue |= (uint32_t)p << (3 * nq); ++nq; }
        if (nq > 0) {
          ++depth;
          continue;
        }
      }
      // showdown with side pots (cfr_gpu.cu:694-752), levels in ascending order
      const int inv_up = G.stack[up] - (int)(sb[up * kWalkWG] & 0xFFFFu);
      float pay = -(float)inv_up;
      int prev = 0;
      for (;;) {
        int lvl = 0x7FFFFFFF;
        for (int p = 0; p < N; ++p) {
          const int inv = G.stack[p] - (int)(sb[p * kWalkWG] & 0xFFFFu);
          if (inv > prev && inv < lvl) lvl = inv;
        }
        if (lvl == 0x7FFFFFFF) break;
        int layer_pot = 0;
        uint32_t best = 0;
        for (int p = 0; p < N; ++p) {
          const int inv = G.stack[p] - (int)(sb[p * kWalkWG] & 0xFFFFu);
          const int c = (inv < lvl ? inv : lvl) - prev;
          layer_pot += c > 0 ? c : 0;
          const uint32_t r = rk[p * kWalkWG];
          if (!(folded >> p & 1u) && inv >= lvl && r > best) best = r;
        }
        if (best > 0) {
          int n_win = 0;
          for (int p = 0; p < N; ++p)
            if (!(folded >> p & 1u) && G.stack[p] - (int)(sb[p * kWalkWG] & 0xFFFFu) >= lvl && rk[p * kWalkWG] == best)
              ++n_win;
          if (!(folded >> up & 1u) && inv_up >= lvl && rk[up * kWalkWG] == best) pay += (float)layer_pot / (float)n_win;
        }
        prev = lvl;
      }
      finish(pay, 2);
      continue;
    }

    // decision node
    const int player = (int)(queue & 7u);
    const uint32_t sbv = sb[player * kWalkWG];
    const int stk = (int)(sbv & 0xFFFFu), bet = (int)(sbv >> 16);
    const int to_call = current_bet - bet;
    const uint32_t vm = valid_mask(G, pot, stk, to_call, current_bet, lfr, raises);
    const uint32_t Bp = bk[player * kWalkWG];
    const uint32_t bbk = street == 0 ? 0u : (Bp >> (8 * street)) & 0xFFu;
    const uint32_t slot = info_hash((uint32_t)player, Bp & 0xFFu, bbk, (uint32_t)street, ab, slot_mask);
    float sg[ROW];
    load_bf16_row<ROW>(T.sig + (size_t)slot * ROW, sg);
    // sample_action (cfr_gpu.cu:510-537) on the BF16 snapshot
    float total = 0.0f;
#pragma unroll
    for (int a = 0; a < ROW; ++a)
      if (a < A && (vm >> a & 1u)) total += sg[a];
    int act = -1;
    if (total < 1e-7f) {
      act = passive_fallback(G, vm);
    } else {
      const float r = u01(rng.next()) * total;
      float cum = 0.0f;
#pragma unroll
      for (int a = 0; a < ROW; ++a) {
        const bool valid = a < A && (vm >> a & 1u);
        if (valid) cum += sg[a];
        if (valid && act < 0 && r <= cum) act = a;
      }
      if (act < 0) act = passive_fallback(G, vm);
    }
    phash = fnv_path(phash, slot, (uint32_t)act);
    ++decisions;
    if (player == up) {
      const uint32_t av = (uint32_t)act | (vm << 8);
      if (npend < kRegPending) {
        ps[npend * kWalkWG] = slot;
        pa[npend * kWalkWG] = av;
      } else {
        T.node_slot[(size_t)npend * L + t] = slot;
        T.node_av[(size_t)npend * L + t] = av;
      }
      ++npend;
      upd_legal += __popc(vm);
    }
    // apply_action_gpu + rebuild_queue (cfr_gpu.cu:543-585)
    queue >>= 3;
    --nq;
    const int chips = chips_lds(G, s_kind, s_num, s_den, act, pot, stk, to_call, current_bet, lfr);
    ab = ((ab << G.hist_bits) | (uint32_t)act) & kHistoryMask;
    if (G.fold_mask >> act & 1u) {
      folded |= 1u << player;
    } else {
      const int ns = stk - chips, nb = bet + chips;
      sb[player * kWalkWG] = (uint32_t)ns | ((uint32_t)nb << 16);
      pot += chips;
      if (ns == 0) all_in |= 1u << player;
      if (nb > current_bet) {
        const int raise = nb - current_bet;
        if (raise >= lfr || current_bet == 0) lfr = raise;
        current_bet = nb;
        ++raises;
        queue = 0;
        nq = 0;
        for (int i = 1, p = wrap(player + 1, N); i < N; ++i, p = wrap(p + 1, N))
          if (!((folded | all_in) >> p & 1u)) { queue |= (uint32_t)p << (3 * nq); ++nq; }
      }
    }
    ++depth;
  }
}

// ---------------------------------------------------------------------------
// k_scatter: contributions of every recorded update-player node
// (cfr_gpu.cu:784-837).  Every node on a path sees the terminal utility u;
// sigma comes from the regret row, which no kernel of the launch has modified.
//
// A workgroup takes kScatterTile traversals, sorts their records by
// (slot, record) in LDS, and sums each slot's run in FP64 in that order, so a
// slot receives one set of global atomics per tile however hot it is (the
// preflop info sets collect tens of thousands of records per launch).  The
// run's slot is appended to the launch's update list.
// ---------------------------------------------------------------------------
// Workgroup sizes the scatter is built for; 4 traversals per thread, sort
// capacity 2 records per traversal per pass.
template <int WG>
__device__ __forceinline__ void block_bitonic_sort(uint64_t* keys, int n_pow2) {
  for (int k = 2; k <= n_pow2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n_pow2; i += WG) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint64_t a = keys[i], b = keys[ixj];
          const bool ascending = (i & k) == 0;
          if ((a > b) == ascending) {
            keys[i] = b;
            keys[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
}

template <int ROW, int kScatterWG>
__global__ __launch_bounds__(kScatterWG) void k_scatter(const DevGame* __restrict__ gp,
                                                        LaunchState* __restrict__ st, DevTables T,
                                                        uint32_t n_lanes) {
  constexpr int kScatterTile = 4 * kScatterWG;  // traversals per workgroup
  constexpr int kScatterCap = 2 * kScatterTile; // records sorted per pass
  __shared__ uint64_t s_keys[kScatterCap];
  __shared__ uint32_t s_run[kScatterCap];
  __shared__ uint32_t s_wsum[kScatterWG / 64];
  __shared__ uint32_t s_runs, s_base;
  const DevGame& G = *gp;
  uint32_t n = (uint32_t)st->n;
  if (n > n_lanes) n = n_lanes;
  const uint32_t t0 = blockIdx.x * kScatterTile;
  if (t0 >= n) return;
  const float tw = (float)st->t_weight;
  const int A = G.num_actions;
  const int tx = threadIdx.x, lane = tx & 63, wave = tx >> 6;

  // records of this thread's traversals: t0 + tx + j*WG, j < 4
  constexpr int kPer = kScatterTile / kScatterWG;
  int cnt[kPer];
  uint32_t mine = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t t = t0 + tx + j * kScatterWG;
    cnt[j] = t < n ? (int)T.node_count[t] : 0;
    mine += (uint32_t)cnt[j];
  }
  // block exclusive scan of the per-thread record counts (waves, then wave totals)
  uint32_t incl = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t v = __shfl_up(incl, d, 64);
    if (lane >= d) incl += v;
  }
  if (lane == 63) s_wsum[wave] = incl;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (int w = 0; w < kScatterWG / 64; ++w) {
    const uint32_t v = s_wsum[w];
    if (w < wave) before += v;
    total += v;
  }
  const uint32_t my_first = before + incl - mine;

  for (uint32_t c0 = 0; c0 < total; c0 += kScatterCap) {
    const uint32_t cn = total - c0 < (uint32_t)kScatterCap ? total - c0 : (uint32_t)kScatterCap;
    int np2 = 64;
    while ((uint32_t)np2 < cn) np2 <<= 1;
    for (int i = tx; i < np2; i += kScatterWG) s_keys[i] = ~0ull;
    if (tx == 0) s_runs = 0;
    __syncthreads();
    // keys: slot << 32 | record reference (local traversal index << 5 | k)
    uint32_t pos = my_first;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const uint32_t tl = (uint32_t)(tx + j * kScatterWG);
      for (int k = 0; k < cnt[j]; ++k, ++pos) {
        if (pos < c0 || pos >= c0 + cn) continue;
        const uint32_t slot = T.node_slot[(size_t)k * T.lanes + t0 + tl];
        s_keys[pos - c0] = ((uint64_t)slot << 32) | (tl << 5) | (uint32_t)k;
      }
    }
    __syncthreads();
    block_bitonic_sort<kScatterWG>(s_keys, np2);
    // each run of equal slots is summed by the thread holding its first record
    for (uint32_t i = tx; i < cn; i += kScatterWG) {
      const uint32_t slot = (uint32_t)(s_keys[i] >> 32);
      if (i > 0 && (uint32_t)(s_keys[i - 1] >> 32) == slot) continue;
      const float* rrow = T.R + (size_t)slot * ROW;
      float rr[ROW];
      load_f32_row<ROW>(rrow, rr);
      double dr[ROW], ds[ROW];
#pragma unroll
      for (int a = 0; a < ROW; ++a) dr[a] = ds[a] = 0.0;
      uint32_t legal_any = 0;
      // sigma of the run's rows, recomputed only when the legal mask changes
      // (hash collisions can put info sets with different masks in one slot)
      uint32_t cached_vm = 0;
      float sig_c[ROW];
      float q_of_act[ROW];
#pragma unroll
      for (int a = 0; a < ROW; ++a) sig_c[a] = q_of_act[a] = 0.0f;
      uint32_t r = i;
      while (r < cn && (uint32_t)(s_keys[r] >> 32) == slot) {
        // gather up to 4 records of the run at once so their loads overlap
        constexpr int kBatch = 4;
        uint32_t av[kBatch];
        float uu[kBatch];
        int m = 0;
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
          av[b] = 0;
          uu[b] = 0.0f;
          if (r + b < cn && (uint32_t)(s_keys[r + b] >> 32) == slot) {
            const uint32_t ref = (uint32_t)s_keys[r + b];
            const uint32_t t = t0 + (ref >> 5), k = ref & 31u;
            av[b] = T.node_av[(size_t)k * T.lanes + t];
            uu[b] = T.node_u[t];
            m = b + 1;
          }
        }
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
          if (b >= m) break;
          const uint32_t act = av[b] & 0xFFu, vm = av[b] >> 8;
          const float u = uu[b];
          if (vm != cached_vm) {
            cached_vm = vm;
            float sum_r = 0.0f;
            float lr[ROW];
#pragma unroll
            for (int a = 0; a < ROW; ++a) {
              lr[a] = (a < A && (vm >> a & 1u)) ? fmaxf(0.0f, rr[a]) : 0.0f;
              sum_r += lr[a];
            }
            const float uni = 1.0f / (float)__popc(vm);
#pragma unroll
            for (int a = 0; a < ROW; ++a) {
              sig_c[a] = (sum_r > 1e-7f) ? lr[a] / sum_r : ((a < A && (vm >> a & 1u)) ? uni : 0.0f);
              q_of_act[a] = fmaxf((sum_r > 1e-7f) ? lr[a] / sum_r : uni, 0.05f);
            }
          }
          legal_any |= vm;
          float q = q_of_act[0];
#pragma unroll
          for (int a = 1; a < ROW; ++a) q = ((uint32_t)a == act) ? q_of_act[a] : q;
          const float rd_sampled = u / q - u;
          const float rd_other = 0.0f - u;
#pragma unroll
          for (int a = 0; a < ROW; ++a)
            if (a < A && (vm >> a & 1u)) {
              dr[a] += (double)((uint32_t)a == act ? rd_sampled : rd_other);
              ds[a] += (double)(tw * sig_c[a]);
            }
        }
        r += (uint32_t)m;
      }
      double* drow = T.D + (size_t)slot * (2 * ROW);
#pragma unroll
      for (int a = 0; a < ROW; ++a)
        if (legal_any >> a & 1u) {
          unsafeAtomicAdd(drow + a, dr[a]);
          unsafeAtomicAdd(drow + ROW + a, ds[a]);
        }
      s_run[atomicAdd(&s_runs, 1u)] = slot;
    }
    __syncthreads();
    if (tx == 0) s_base = atomicAdd(&st->runs, s_runs);
    __syncthreads();
    for (uint32_t i = tx; i < s_runs; i += kScatterWG) T.upd_list[s_base + i] = s_run[i];
    __syncthreads();  // s_keys and s_run are reused by the next chunk
  }
}
// ---------------------------------------------------------------------------
// k_update: the scatter's run list holds every touched slot (once per tile);
// the epoch stamp lets exactly one thread handle each slot.
// R <- clamp(R + dR), S <- S + dS, sigma <- RM(R), deltas cleared.
// ---------------------------------------------------------------------------
template <int ROW>
__global__ __launch_bounds__(256) void k_update(const DevGame* __restrict__ gp, const LaunchState* __restrict__ st,
                                                DevTables T) {
  const DevGame& G = *gp;
  const bool cfr_plus = G.use_cfr_plus != 0;
  const uint32_t epoch = st->epoch;
  const uint32_t runs = st->runs;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < runs; i += gridDim.x * blockDim.x) {
    const uint32_t slot = T.upd_list[i];
    if (atomicExch(T.stamp + slot, epoch) == epoch) continue;
    float* R = T.R + (size_t)slot * ROW;
    float* S = T.S + (size_t)slot * ROW;
    double* D = T.D + (size_t)slot * (2 * ROW);
    float r[ROW], s[ROW];
    load_f32_row<ROW>(R, r);
    load_f32_row<ROW>(S, s);
    float sum = 0.0f;
#pragma unroll
    for (int a = 0; a < ROW; ++a) {
      float nr = (float)((double)r[a] + D[a]);
      if (cfr_plus) nr = fmaxf(0.0f, nr);
      r[a] = nr;
      s[a] = (float)((double)s[a] + D[ROW + a]);
      sum += fmaxf(0.0f, nr);
    }
    uint32_t packed[ROW / 2];
#pragma unroll
    for (int a = 0; a < ROW; a += 2) {
      const float x0 = (sum > 1e-7f) ? fmaxf(0.0f, r[a]) / sum : 0.0f;
      const float x1 = (sum > 1e-7f) ? fmaxf(0.0f, r[a + 1]) / sum : 0.0f;
      packed[a / 2] = (uint32_t)f32_to_bf16(x0) | ((uint32_t)f32_to_bf16(x1) << 16);
    }
#pragma unroll
    for (int v = 0; v < ROW / 4; ++v) {
      reinterpret_cast<float4*>(R)[v] = make_float4(r[4 * v], r[4 * v + 1], r[4 * v + 2], r[4 * v + 3]);
      reinterpret_cast<float4*>(S)[v] = make_float4(s[4 * v], s[4 * v + 1], s[4 * v + 2], s[4 * v + 3]);
    }
#pragma unroll
    for (int v = 0; v < ROW / 8; ++v)
      reinterpret_cast<uint4*>(T.sig + (size_t)slot * ROW)[v] =
          make_uint4(packed[4 * v], packed[4 * v + 1], packed[4 * v + 2], packed[4 * v + 3]);
#pragma unroll
    for (int v = 0; v < ROW / 2; ++v) reinterpret_cast<double4*>(D)[v] = make_double4(0.0, 0.0, 0.0, 0.0);
  }
}

}  // namespace

void launch_traverse(hipStream_t s, int row, const DevGame* g, LaunchState* st, BucketTables bt,
                     ArithTables at, const uint8_t* lut, DevTables t, uint32_t n_lanes, uint32_t walk_blocks,
                     uint32_t scatter_wg, mccfr_trace* trace) {
  hipLaunchKernelGGL(k_deal, dim3(grid_for(n_lanes, kDealWG, 1u << 30)), dim3(kDealWG), 0, s, g, st, bt, at, lut, t,
                     n_lanes);
  const dim3 wgrid(grid_for(n_lanes, kWalkWG, walk_blocks ? walk_blocks : 1u << 30));
  if (row == 8)
    hipLaunchKernelGGL(k_walk<8>, wgrid, dim3(kWalkWG), 0, s, g, st, t, n_lanes, trace);
  else
    hipLaunchKernelGGL(k_walk<16>, wgrid, dim3(kWalkWG), 0, s, g, st, t, n_lanes, trace);
  const unsigned wg = scatter_wg == 256 || scatter_wg == 512 ? scatter_wg : 1024;
  const dim3 sgrid(grid_for(n_lanes, 4 * wg, 1u << 30));
#define MCCFR_SCATTER(R, W) hipLaunchKernelGGL((k_scatter<R, W>), sgrid, dim3(W), 0, s, g, st, t, n_lanes)
  if (row == 8) {
    if (wg == 256) MCCFR_SCATTER(8, 256);
    else if (wg == 512) MCCFR_SCATTER(8, 512);
    else MCCFR_SCATTER(8, 1024);
  } else {
    if (wg == 256) MCCFR_SCATTER(16, 256);
    else if (wg == 512) MCCFR_SCATTER(16, 512);
    else MCCFR_SCATTER(16, 1024);
  }
#undef MCCFR_SCATTER
}

void launch_update(hipStream_t s, int row, const DevGame* g, LaunchState* st, DevTables t, uint32_t n_lanes) {
  const unsigned grid = grid_for((size_t)n_lanes * 2, 256, 8192);
  if (row == 8)
    hipLaunchKernelGGL(k_update<8>, dim3(grid), dim3(256), 0, s, g, st, t);
  else
    hipLaunchKernelGGL(k_update<16>, dim3(grid), dim3(256), 0, s, g, st, t);
}

}  // namespace mccfr
