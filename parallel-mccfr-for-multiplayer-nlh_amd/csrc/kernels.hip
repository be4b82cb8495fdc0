// gfx950 kernels of the MCCFR hot path.
//
//   k_traverse<ROW>   one traversal per lane: deal, buckets, betting-tree walk,
//                     delta scatter (reference semantics of kernel_simulate_batch,
//                     cfr_gpu.cu:878-1011, made iterative: both node types follow
//                     exactly one child, cfr_gpu.cu:800 and :844, so a traversal
//                     is a single root-to-leaf path).
//   k_update<ROW>     touched-slot regret/strategy-sum update, CFR+ clamp and
//                     regret matching of the next launch's BF16 snapshot
//                     (kernel_clamp_regrets + kernel_regret_matching,
//                     cfr_gpu.cu:1018-1022, :318-351, restricted to the slots
//                     the launch touched: both are per-slot and idempotent on
//                     untouched slots).
//   k_rm_sweep<ROW>   full-table regret matching (after tables are uploaded).
//   k_normalize<ROW>  average strategy (kernel_normalize_strategy, cfr_gpu.cu:1029-1058).
//   k_eval7_walk      7-card rank via the rank-walk table (eval7_gpu, cfr_gpu.cu:126-137).
//   k_eval7_arith     7-card rank via flush table + perfect hash of the rank multiset.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mccfr {
namespace {

constexpr int kWG = 256;             // traversal workgroup (4 wave64)
constexpr int kDeckStride = 52;      // bytes of LDS deck per lane (13 dwords: odd stride, no bank aliasing)
constexpr int kMaxPending = kMaxDepth;

template <class T>
__device__ __forceinline__ T sel6(const T (&a)[kMaxPlayers], int i) {
  T v = a[0];
#pragma unroll
  for (int k = 1; k < kMaxPlayers; ++k) v = (i == k) ? a[k] : v;
  return v;
}
template <class T>
__device__ __forceinline__ void add6(T (&a)[kMaxPlayers], int i, T d) {
#pragma unroll
  for (int k = 0; k < kMaxPlayers; ++k) a[k] += (i == k) ? d : (T)0;
}
template <int ROW>
__device__ __forceinline__ float selrow(const float (&a)[ROW], int i) {
  float v = a[0];
#pragma unroll
  for (int k = 1; k < ROW; ++k) v = (i == k) ? a[k] : v;
  return v;
}

template <int ROW>
__device__ __forceinline__ void load_bf16_row(const uint16_t* p, float (&out)[ROW]) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
  for (int v = 0; v < ROW / 8; ++v) {
    const uint4 w = q[v];
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      out[v * 8 + 2 * k] = bf16_to_f32(ws[k] & 0xFFFFu);
      out[v * 8 + 2 * k + 1] = bf16_to_f32(ws[k] >> 16);
    }
  }
}
template <int ROW>
__device__ __forceinline__ void load_f32_row(const float* p, float (&out)[ROW]) {
  const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
  for (int v = 0; v < ROW / 4; ++v) {
    const float4 w = q[v];
    out[4 * v] = w.x;
    out[4 * v + 1] = w.y;
    out[4 * v + 2] = w.z;
    out[4 * v + 3] = w.w;
  }
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

__device__ __forceinline__ void touch_slot(const DevTables& T, LaunchState* st, uint32_t slot,
                                           uint32_t epoch) {
  if (atomicExch(T.stamp + slot, epoch) != epoch) {
    const uint32_t i = atomicAdd(&st->touched, 1u);
    if (i < T.touched_cap) T.touched[i] = slot;
    else atomicOr(&st->overflow, 1u);
  }
}

__device__ __forceinline__ uint32_t fnv_path(uint32_t h, uint32_t slot, uint32_t a) {
  h = (h ^ (slot & 0xFFu)) * 16777619u;
  h = (h ^ ((slot >> 8) & 0xFFu)) * 16777619u;
  h = (h ^ ((slot >> 16) & 0xFFu)) * 16777619u;
  h = (h ^ (slot >> 24)) * 16777619u;
  return (h ^ (a & 0xFFu)) * 16777619u;
}

// ---------------------------------------------------------------------------
// Traversal kernel.
// ---------------------------------------------------------------------------
template <int ROW>
__global__ __launch_bounds__(kWG) void k_traverse(const DevGame* __restrict__ gp,
                                                  LaunchState* __restrict__ st,
                                                  const int32_t* __restrict__ hr,
                                                  const uint8_t* __restrict__ lut_g, DevTables T,
                                                  uint32_t n_lanes, mccfr_trace* __restrict__ trace) {
  __shared__ uint32_t s_lut[2704 / 4];
  __shared__ uint8_t s_deck[kDeckStride * kWG];
  __shared__ uint8_t s_kind[MCCFR_MAX_ACTIONS], s_num[MCCFR_MAX_ACTIONS], s_den[MCCFR_MAX_ACTIONS];
  const DevGame& G = *gp;
  for (int i = threadIdx.x; i < 2704 / 4; i += kWG) s_lut[i] = reinterpret_cast<const uint32_t*>(lut_g)[i];
  if (threadIdx.x < MCCFR_MAX_ACTIONS) {
    s_kind[threadIdx.x] = G.kind[threadIdx.x];
    s_num[threadIdx.x] = G.num[threadIdx.x];
    s_den[threadIdx.x] = G.den[threadIdx.x];
  }
  __syncthreads();

  const uint32_t lane = blockIdx.x * kWG + threadIdx.x;
  const uint32_t n = (uint32_t)st->n;
  if (lane >= n || lane >= n_lanes) return;
  const int up = st->update_player;
  const float tw = (float)st->t_weight;
  const uint32_t epoch = st->epoch;
  const int N = G.N;
  const int A = G.num_actions;
  const uint32_t slot_mask = (uint32_t)((1ull << G.table_log2) - 1ull);

  // ---- deal: partial Fisher-Yates over 2N+5 draws (cfr_gpu.cu:460-470, 931-937)
  DrawStream rng;
  rng.init(G.seed_lo, G.seed_hi, st->first_tid + lane, G.tag);
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

  // ---- buckets.  The rank-walk state depends only on the set of cards seen,
  // so the board's 3-card prefix is walked once and each player adds hole
  // cards to it: the 5-card and 6-card words equal the reference's eval5/eval6
  // reads (cfr_gpu.cu:114-150) and the 7-card word its eval7, in 3 + 4N
  // gathers instead of 18N.
  const uint32_t c0 = dk[2 * N], c1 = dk[2 * N + 1], c2 = dk[2 * N + 2];
  const int rm3 = card_rm(dk[2 * N + 3]), rm4 = card_rm(dk[2 * N + 4]);
  int pw = hr[53 + card_rm(c0)];
  pw = hr[pw + card_rm(c1)];
  const int p3 = hr[pw + card_rm(c2)];
  uint32_t B[kMaxPlayers], RK[kMaxPlayers];
  const uint8_t* lut = reinterpret_cast<const uint8_t*>(s_lut);
#pragma unroll
  for (int pl = 0; pl < kMaxPlayers; ++pl) {
    B[pl] = 0;
    RK[pl] = 0;
    if (pl < N) {
      const uint32_t h0 = dk[2 * pl], h1 = dk[2 * pl + 1];
      const uint32_t hb = lut[h0 * 52 + h1];
      const int q4 = hr[p3 + card_rm(h0)];
      const int p5 = hr[q4 + card_rm(h1)];
      const int p6 = hr[p5 + rm3];
      const uint32_t r7 = decode_packed((uint32_t)hr[p6 + rm4] & 0xFFFFu);
      const uint32_t bf = bucket_of_rank(decode_packed((uint32_t)p5 & 0xFFFFu));
      const uint32_t bt = bucket_of_rank(decode_packed((uint32_t)p6 & 0xFFFFu));
      B[pl] = hb | (bf << 8) | (bt << 16) | (bucket_of_rank(r7) << 24);
      RK[pl] = r7;
    }
  }

  // ---- blinds and preflop queue (cfr_gpu.cu:954-981)
  int stacks[kMaxPlayers], bets[kMaxPlayers], start[kMaxPlayers];
#pragma unroll
  for (int pl = 0; pl < kMaxPlayers; ++pl) {
    start[pl] = pl < N ? G.stack[pl] : 0;
    stacks[pl] = start[pl];
    bets[pl] = 0;
  }
  uint32_t folded = 0, all_in = 0;
  int pot, current_bet, lfr = G.bb, street = 0, raises = 0;
  uint32_t ab = 0;
  const int dealer = (int)(rng.next() % (uint32_t)N);
  const int sb_pos = (N == 2) ? dealer : (dealer + 1) % N;
  const int bb_pos = (N == 2) ? (dealer + 1) % N : (dealer + 2) % N;
  {
    const int s_sb = sel6(stacks, sb_pos);
    const int sb_amt = G.sb < s_sb ? G.sb : s_sb;
    add6(stacks, sb_pos, -sb_amt);
    add6(bets, sb_pos, sb_amt);
    const int s_bb = sel6(stacks, bb_pos);
    const int bb_amt = G.bb < s_bb ? G.bb : s_bb;
    add6(stacks, bb_pos, -bb_amt);
    add6(bets, bb_pos, bb_amt);
    pot = sb_amt + bb_amt;
    current_bet = bb_amt;
    if (sel6(stacks, sb_pos) == 0) all_in |= 1u << sb_pos;
    if (sel6(stacks, bb_pos) == 0) all_in |= 1u << bb_pos;
  }
  uint32_t queue = 0;
  int nq = 0;
  for (int i = 0; i < N; ++i) {
    const int p = (bb_pos + 1 + i) % N;
    if (!(all_in >> p & 1u)) { queue |= (uint32_t)p << (3 * nq); ++nq; }
  }

  // ---- the path (cfr_gpu.cu:644-860 as a loop)
  const uint32_t allmask = (1u << N) - 1u;
  float u = 0.0f;
  int term = 0, decisions = 0, npend = 0, upd_legal = 0;
  uint32_t phash = 2166136261u;
  uint32_t pend_slot[kMaxPending], pend_av[kMaxPending];
  float pend_q[kMaxPending];
  for (int depth = 0;;) {
    if (depth >= kMaxDepth) { term = 3; u = 0.0f; break; }
    const uint32_t live = allmask & ~folded;
    if (__popc(live) == 1) {
      term = 1;
      const int inv_up = sel6(start, up) - sel6(stacks, up);
      u = ((int)__builtin_ctz(live) == up) ? (float)(pot - inv_up) : (float)(-inv_up);
      break;
    }
    if (nq == 0) {
      if (street < 3) {
        ++street;
#pragma unroll
        for (int pl = 0; pl < kMaxPlayers; ++pl) bets[pl] = 0;
        current_bet = 0;
        lfr = G.bb;
        raises = 0;
        for (int i = 1; i <= N; ++i) {
          const int p = (dealer + i) % N;
          if (!((folded | all_in) >> p & 1u)) { queue |= (uint32_t)p << (3 * nq); ++nq; }
        }
        if (nq > 0) { ++depth; continue; }
      }
      // showdown with side pots (cfr_gpu.cu:694-752), levels in ascending order
      term = 2;
      const int inv_up = sel6(start, up) - sel6(stacks, up);
      float pay = -(float)inv_up;
      int prev = 0;
      for (;;) {
        int L = 0x7FFFFFFF;
#pragma unroll
        for (int pl = 0; pl < kMaxPlayers; ++pl) {
          const int inv = start[pl] - stacks[pl];
          if (pl < N && inv > prev && inv < L) L = inv;
        }
        if (L == 0x7FFFFFFF) break;
        int layer_pot = 0;
        uint32_t best = 0;
#pragma unroll
        for (int pl = 0; pl < kMaxPlayers; ++pl) {
          if (pl >= N) continue;
          const int inv = start[pl] - stacks[pl];
          const int c = (inv < L ? inv : L) - prev;
          layer_pot += c > 0 ? c : 0;
          if (!(folded >> pl & 1u) && inv >= L && RK[pl] > best) best = RK[pl];
        }
        if (best > 0) {
          int n_win = 0;
#pragma unroll
          for (int pl = 0; pl < kMaxPlayers; ++pl)
            if (pl < N && !(folded >> pl & 1u) && start[pl] - stacks[pl] >= L && RK[pl] == best) ++n_win;
          if (!(folded >> up & 1u) && inv_up >= L && sel6(RK, up) == best)
            pay += (float)layer_pot / (float)n_win;
        }
        prev = L;
      }
      u = pay;
      break;
    }

    // decision node
    const int player = (int)(queue & 7u);
    const int stk = sel6(stacks, player), bet = sel6(bets, player);
    const int to_call = current_bet - bet;
    const uint32_t vm = valid_mask(G, pot, stk, to_call, current_bet, lfr, raises);
    const uint32_t Bp = sel6(B, player);
    const uint32_t bbk = street == 0 ? 0u : (Bp >> (8 * street)) & 0xFFu;
    const uint32_t slot = info_hash((uint32_t)player, Bp & 0xFFu, bbk, (uint32_t)street, ab, slot_mask);
    float sg[ROW];
    load_bf16_row<ROW>(T.sig + (size_t)slot * ROW, sg);
    const bool is_up = player == up;
    float strat[ROW];
    if (is_up) {
      float rr[ROW];
      load_f32_row<ROW>(T.R + (size_t)slot * ROW, rr);
      float sum_r = 0.0f;
#pragma unroll
      for (int a = 0; a < ROW; ++a) {
        const float r = (a < A && (vm >> a & 1u)) ? fmaxf(0.0f, rr[a]) : 0.0f;
        strat[a] = r;
        sum_r += r;
      }
      const float uni = 1.0f / (float)__popc(vm);
#pragma unroll
      for (int a = 0; a < ROW; ++a)
        strat[a] = (sum_r > 1e-7f) ? strat[a] / sum_r : ((a < A && (vm >> a & 1u)) ? uni : 0.0f);
    }
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
    if (is_up) {
      const float q = fmaxf(selrow<ROW>(strat, act), 0.05f);
      double* drow = T.D + (size_t)slot * (2 * ROW);
#pragma unroll
      for (int a = 0; a < ROW; ++a)
        if (a < A && (vm >> a & 1u)) unsafeAtomicAdd(drow + ROW + a, (double)(tw * strat[a]));
      touch_slot(T, st, slot, epoch);
      upd_legal += __popc(vm);
      pend_slot[npend] = slot;
      pend_q[npend] = q;
      pend_av[npend] = (uint32_t)act | (vm << 8);
      ++npend;
    }
    // apply_action_gpu + rebuild_queue (cfr_gpu.cu:543-585)
    queue >>= 3;
    --nq;
    const int chips = chips_lds(G, s_kind, s_num, s_den, act, pot, stk, to_call, current_bet, lfr);
    ab = ((ab << G.hist_bits) | (uint32_t)act) & kHistoryMask;
    if (G.fold_mask >> act & 1u) {
      folded |= 1u << player;
    } else {
      add6(stacks, player, -chips);
      add6(bets, player, chips);
      pot += chips;
      if (stk - chips == 0) all_in |= 1u << player;
      const int nb = bet + chips;
      if (nb > current_bet) {
        const int raise = nb - current_bet;
        if (raise >= lfr || current_bet == 0) lfr = raise;
        current_bet = nb;
        ++raises;
        queue = 0;
        nq = 0;
        for (int i = 1; i < N; ++i) {
          const int p = (player + i) % N;
          if (!((folded | all_in) >> p & 1u)) { queue |= (uint32_t)p << (3 * nq); ++nq; }
        }
      }
    }
    ++depth;
  }

  // regret deltas of the update nodes on the path, now that u is known (cfr_gpu.cu:825-837)
  for (int k = 0; k < npend; ++k) {
    const uint32_t slot = pend_slot[k];
    const uint32_t act = pend_av[k] & 0xFFu, vm = pend_av[k] >> 8;
    const float cf = u / pend_q[k];
    const float rd_sampled = cf - u;
    const float rd_other = 0.0f - u;
    double* drow = T.D + (size_t)slot * (2 * ROW);
#pragma unroll
    for (int a = 0; a < ROW; ++a)
      if (a < A && (vm >> a & 1u)) unsafeAtomicAdd(drow + a, (double)((uint32_t)a == act ? rd_sampled : rd_other));
  }
  if (trace) {
    mccfr_trace t;
    t.utility = u;
    t.path_hash = phash;
    t.decisions = (uint16_t)decisions;
    t.updates = (uint8_t)npend;
    t.terminal = (uint8_t)term;
    t.update_legal = (uint16_t)upd_legal;
    t.reserved = 0;
    trace[lane] = t;
  }
}

// ---------------------------------------------------------------------------
// Touched-slot update: R <- clamp(R + dR), S <- S + dS, sigma <- RM(R), deltas cleared.
// ---------------------------------------------------------------------------
template <int ROW>
__global__ __launch_bounds__(256) void k_update(const DevGame* __restrict__ gp,
                                                const LaunchState* __restrict__ st, DevTables T) {
  const DevGame& G = *gp;
  uint32_t count = st->touched;
  if (count > T.touched_cap) count = T.touched_cap;
  const bool cfr_plus = G.use_cfr_plus != 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const uint32_t slot = T.touched[i];
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

template <int ROW>
__global__ __launch_bounds__(256) void k_rm_sweep(DevTables T, uint64_t slots) {
  for (uint64_t slot = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; slot < slots;
       slot += (uint64_t)gridDim.x * blockDim.x) {
    float r[ROW];
    load_f32_row<ROW>(T.R + slot * ROW, r);
    float sum = 0.0f;
#pragma unroll
    for (int a = 0; a < ROW; ++a) sum += fmaxf(0.0f, r[a]);
    uint32_t packed[ROW / 2];
#pragma unroll
    for (int a = 0; a < ROW; a += 2) {
      const float x0 = (sum > 1e-7f) ? fmaxf(0.0f, r[a]) / sum : 0.0f;
      const float x1 = (sum > 1e-7f) ? fmaxf(0.0f, r[a + 1]) / sum : 0.0f;
      packed[a / 2] = (uint32_t)f32_to_bf16(x0) | ((uint32_t)f32_to_bf16(x1) << 16);
    }
#pragma unroll
    for (int v = 0; v < ROW / 8; ++v)
      reinterpret_cast<uint4*>(T.sig + slot * ROW)[v] =
          make_uint4(packed[4 * v], packed[4 * v + 1], packed[4 * v + 2], packed[4 * v + 3]);
  }
}

template <int ROW>
__global__ __launch_bounds__(256) void k_normalize(const float* __restrict__ S, float* __restrict__ out,
                                                   uint64_t slots) {
  for (uint64_t slot = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; slot < slots;
       slot += (uint64_t)gridDim.x * blockDim.x) {
    float s[ROW];
    load_f32_row<ROW>(S + slot * ROW, s);
    float total = 0.0f;
#pragma unroll
    for (int a = 0; a < ROW; ++a) total += s[a];
#pragma unroll
    for (int v = 0; v < ROW / 4; ++v) {
      float4 o;
      o.x = (total > 1e-7f) ? s[4 * v] / total : 0.0f;
      o.y = (total > 1e-7f) ? s[4 * v + 1] / total : 0.0f;
      o.z = (total > 1e-7f) ? s[4 * v + 2] / total : 0.0f;
      o.w = (total > 1e-7f) ? s[4 * v + 3] / total : 0.0f;
      reinterpret_cast<float4*>(out + slot * ROW)[v] = o;
    }
  }
}

__global__ void k_set_state(LaunchState* st, int up, int64_t iteration, int64_t t_weight,
                            uint64_t first_tid, uint32_t n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->update_player = up;
    st->iteration = iteration;
    st->t_weight = t_weight;
    st->first_tid = first_tid;
    st->n = (int32_t)n;
    uint32_t e = st->epoch + 1u;
    st->epoch = e ? e : 1u;
    st->touched = 0;
  }
}

// Advances the state to the next launch of the training loop
// (cfr_gpu.cu:1303-1318: player p+1, or player 0 of iteration+1).
__global__ void k_advance_state(LaunchState* st, int num_players, int use_linear_cfr, uint32_t batch) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->first_tid += batch;
    st->n = (int32_t)batch;
    int p = st->update_player + 1;
    if (p >= num_players) {
      p = 0;
      st->iteration += 1;
    }
    st->update_player = p;
    st->t_weight = use_linear_cfr ? st->iteration : 1;
    uint32_t e = st->epoch + 1u;
    st->epoch = e ? e : 1u;
    st->touched = 0;
  }
}

// ---------------------------------------------------------------------------
// Seven-card evaluators.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_hand(const uint8_t* cards, size_t stride, size_t i, uint32_t c[7]) {
  const uint8_t* p = cards + i * stride;
  if (stride == 8) {
    const uint2 w = *reinterpret_cast<const uint2*>(p);
    c[0] = w.x & 0xFF; c[1] = (w.x >> 8) & 0xFF; c[2] = (w.x >> 16) & 0xFF; c[3] = w.x >> 24;
    c[4] = w.y & 0xFF; c[5] = (w.y >> 8) & 0xFF; c[6] = (w.y >> 16) & 0xFF;
  } else {
#pragma unroll
    for (int k = 0; k < 7; ++k) c[k] = p[k];
  }
}

__device__ __forceinline__ bool distinct_cards(const uint32_t c[7]) {
  uint64_t m = 0;
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    ok = ok && c[k] < 52u;
    m |= 1ull << (c[k] & 63u);
  }
  return ok && __popcll(m) == 7;
}

__global__ __launch_bounds__(256) void k_eval7_walk(const int32_t* __restrict__ hr,
                                                    const uint8_t* __restrict__ cards, size_t stride,
                                                    uint16_t* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    uint32_t c[7];
    load_hand(cards, stride, i, c);
    if (!distinct_cards(c)) {  // codes above 51 or repeated cards have no rank
      out[i] = 0;
      continue;
    }
    int p = hr[53 + card_rm(c[0])];
#pragma unroll
    for (int k = 1; k < 6; ++k) p = hr[p + card_rm(c[k])];
    out[i] = (uint16_t)decode_packed((uint32_t)hr[p + card_rm(c[6])] & 0xFFFFu);
  }
}

// Arithmetic evaluator.  A suit holding >= 5 of the 7 cards makes the hand a
// flush or straight flush (no full house or quads fits beside five suited
// cards among seven), ranked by a 13-bit rank-mask table.  Otherwise the rank
// multiset (counts c_0..c_12, sum 7, each <= 4) is ranked lexicographically
// by a minimal perfect hash: index = sum_i off[i][k_i][c_i] with k_i the cards
// left before rank i; the 13 lookups are independent because k_i comes from a
// prefix sum of 3-bit count fields done by one multiply.
constexpr int kArithWG = 512;
__global__ __launch_bounds__(kArithWG) void k_eval7_arith(ArithTables t, const uint8_t* __restrict__ cards,
                                                          size_t stride, uint16_t* __restrict__ out, size_t n) {
  __shared__ uint16_t s_nf[kNonFlushClasses + 1];
  __shared__ uint16_t s_fl[8192];
  __shared__ uint16_t s_off[13 * 8 * 5];
  for (int i = threadIdx.x; i < kNonFlushClasses; i += kArithWG) s_nf[i] = t.nonflush[i];
  for (int i = threadIdx.x; i < 8192; i += kArithWG) s_fl[i] = t.flush[i];
  for (int i = threadIdx.x; i < 13 * 8 * 5; i += kArithWG) s_off[i] = t.offsets[i];
  __syncthreads();
  for (size_t i = (size_t)blockIdx.x * kArithWG + threadIdx.x; i < n; i += (size_t)gridDim.x * kArithWG) {
    uint32_t c[7];
    load_hand(cards, stride, i, c);
    if (!distinct_cards(c)) {
      out[i] = 0;
      continue;
    }
    uint64_t cnt = 0, cmask = 0;
    uint32_t scnt = 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const uint32_t s = (c[k] * 79u) >> 10;  // c / 13 for c < 52
      const uint32_t r = c[k] - 13u * s;
      cnt += 1ull << (3 * r);
      scnt += 1u << (4 * s);
      cmask |= 1ull << c[k];
    }
    const uint32_t fl = (scnt + 0x3333u) & 0x8888u;  // nibble >= 5
    uint32_t rank;
    if (fl) {
      const uint32_t s = (uint32_t)__builtin_ctz(fl) >> 2;
      rank = s_fl[(uint32_t)(cmask >> (13 * s)) & 0x1FFFu];
    } else {
      const uint64_t pre = cnt * 0x0000001249249249ull;  // field i = c_0 + .. + c_i
      uint32_t idx = 0;
#pragma unroll
      for (int r = 0; r < 13; ++r) {
        const uint32_t before = r == 0 ? 0u : (uint32_t)(pre >> (3 * (r - 1))) & 7u;
        const uint32_t cr = (uint32_t)(cnt >> (3 * r)) & 7u;
        idx += s_off[(r * 8 + (7 - before)) * 5 + cr];
      }
      rank = s_nf[idx];
    }
    out[i] = (uint16_t)rank;
  }
}

// Hands of 7 distinct uniform cards: the first 7 positions of a Fisher-Yates
// shuffle driven by the traversal draw stream of id i.
__global__ __launch_bounds__(256) void k_random_hands(uint32_t seed_lo, uint32_t seed_hi, uint32_t tag,
                                                      uint8_t* __restrict__ out, size_t n) {
  __shared__ uint8_t s_deck[kDeckStride * 256];
  uint8_t* dk = s_deck + threadIdx.x * kDeckStride;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    DrawStream rng;
    rng.init(seed_lo, seed_hi, i, tag);
    for (int k = 0; k < 52; ++k) dk[k] = (uint8_t)k;
    for (int k = 0; k < 7; ++k) {
      const uint32_t r = rng.next();
      const int j = k + (int)(((uint64_t)r * (uint64_t)(52 - k)) >> 32);
      const uint8_t a = dk[k], b = dk[j];
      dk[k] = b;
      dk[j] = a;
    }
    uint2 w;
    w.x = (uint32_t)dk[0] | ((uint32_t)dk[1] << 8) | ((uint32_t)dk[2] << 16) | ((uint32_t)dk[3] << 24);
    w.y = (uint32_t)dk[4] | ((uint32_t)dk[5] << 8) | ((uint32_t)dk[6] << 16);
    *reinterpret_cast<uint2*>(out + i * 8) = w;
  }
}

template <int ROW>
__global__ __launch_bounds__(256) void k_randomize(DevTables T, uint64_t slots, int num_actions, uint32_t seed_lo,
                                                   uint32_t seed_hi) {
  for (uint64_t slot = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; slot < slots;
       slot += (uint64_t)gridDim.x * blockDim.x) {
    DrawStream rng;
    rng.init(seed_lo, seed_hi, slot, 0x72616E64u);
    float r[ROW];
#pragma unroll
    for (int a = 0; a < ROW; ++a) r[a] = a < num_actions ? u01(rng.next()) : 0.0f;
#pragma unroll
    for (int v = 0; v < ROW / 4; ++v) {
      reinterpret_cast<float4*>(T.R + slot * ROW)[v] = make_float4(r[4 * v], r[4 * v + 1], r[4 * v + 2], r[4 * v + 3]);
      reinterpret_cast<float4*>(T.S + slot * ROW)[v] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}

__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

inline unsigned grid_for(size_t n, unsigned block, unsigned cap) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (unsigned)(g < cap ? g : cap);
}

}  // namespace

void launch_set_state(hipStream_t s, LaunchState* st, int up, int64_t iteration, int64_t t_weight,
                      uint64_t first_tid, uint32_t n) {
  hipLaunchKernelGGL(k_set_state, dim3(1), dim3(64), 0, s, st, up, iteration, t_weight, first_tid, n);
}
void launch_advance_state(hipStream_t s, LaunchState* st, int num_players, int use_linear_cfr, uint32_t batch) {
  hipLaunchKernelGGL(k_advance_state, dim3(1), dim3(64), 0, s, st, num_players, use_linear_cfr, batch);
}
void launch_traverse(hipStream_t s, int row, const DevGame* g, LaunchState* st, const int32_t* hr,
                     const uint8_t* lut, DevTables t, uint32_t n_lanes, mccfr_trace* trace) {
  const dim3 grid((n_lanes + kWG - 1) / kWG);
  if (row == 8)
    hipLaunchKernelGGL(k_traverse<8>, grid, dim3(kWG), 0, s, g, st, hr, lut, t, n_lanes, trace);
  else
    hipLaunchKernelGGL(k_traverse<16>, grid, dim3(kWG), 0, s, g, st, hr, lut, t, n_lanes, trace);
}
void launch_update(hipStream_t s, int row, const DevGame* g, LaunchState* st, DevTables t) {
  const unsigned grid = grid_for(t.touched_cap, 256, 8192);
  if (row == 8)
    hipLaunchKernelGGL(k_update<8>, dim3(grid), dim3(256), 0, s, g, st, t);
  else
    hipLaunchKernelGGL(k_update<16>, dim3(grid), dim3(256), 0, s, g, st, t);
}
void launch_rm_sweep(hipStream_t s, int row, DevTables t, uint64_t slots) {
  const unsigned grid = grid_for(slots, 256, 65536);
  if (row == 8)
    hipLaunchKernelGGL(k_rm_sweep<8>, dim3(grid), dim3(256), 0, s, t, slots);
  else
    hipLaunchKernelGGL(k_rm_sweep<16>, dim3(grid), dim3(256), 0, s, t, slots);
}
void launch_normalize(hipStream_t s, int row, const float* S, float* out, uint64_t slots) {
  const unsigned grid = grid_for(slots, 256, 65536);
  if (row == 8)
    hipLaunchKernelGGL(k_normalize<8>, dim3(grid), dim3(256), 0, s, S, out, slots);
  else
    hipLaunchKernelGGL(k_normalize<16>, dim3(grid), dim3(256), 0, s, S, out, slots);
}
void launch_eval7_walk(hipStream_t s, const int32_t* hr, const uint8_t* cards, size_t stride, uint16_t* out,
                       size_t n) {
  hipLaunchKernelGGL(k_eval7_walk, dim3(grid_for(n, 256, 1u << 20)), dim3(256), 0, s, hr, cards, stride, out, n);
}
void launch_eval7_arith(hipStream_t s, ArithTables t, const uint8_t* cards, size_t stride, uint16_t* out,
                        size_t n) {
  hipLaunchKernelGGL(k_eval7_arith, dim3(grid_for(n, kArithWG, 2048)), dim3(kArithWG), 0, s, t, cards, stride,
                     out, n);
}
void launch_random_hands(hipStream_t s, uint64_t seed, uint32_t tag, uint8_t* out, size_t n) {
  hipLaunchKernelGGL(k_random_hands, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, s, (uint32_t)seed,
                     (uint32_t)(seed >> 32), tag, out, n);
}
void launch_randomize(hipStream_t s, int row, int num_actions, DevTables t, uint64_t slots, uint64_t seed) {
  const unsigned grid = grid_for(slots, 256, 65536);
  if (row == 8)
    hipLaunchKernelGGL(k_randomize<8>, dim3(grid), dim3(256), 0, s, t, slots, num_actions, (uint32_t)seed,
                       (uint32_t)(seed >> 32));
  else
    hipLaunchKernelGGL(k_randomize<16>, dim3(grid), dim3(256), 0, s, t, slots, num_actions, (uint32_t)seed,
                       (uint32_t)(seed >> 32));
}
void launch_copy(hipStream_t s, const uint4* src, uint4* dst, size_t n16) {
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n16, 256, 1u << 16)), dim3(256), 0, s, src, dst, n16);
}

}  // namespace mccfr
