// evaluate.hip — the strategy-quality harness on the device.
//
// evaluate_strategy_np (poker_cuda/src/eval.cpp:581-627) plays n hands of
// no-limit hold'em: seat 0 follows the normalised average strategy, every
// other seat a scripted archetype (eval.cpp:100-157), the dealer button at
// hand % n.  The reference plays the hands one after another on one
// std::mt19937; here every hand is one lane with its own counter-based stream
// (Philox(counter = {k>>2, hand_lo, hand_hi, 0xE7A1}, key = seed): 2n+5
// partial Fisher-Yates draws on a fresh deck, then uniforms (x>>8) * 2^-24 and
// RANDOM picks x % count in decision order), so any hand can be replayed alone
// and a launch of 2^20 hands fills the chip.
//
// Per-lane game state that is indexed by seat (stacks, bets, investments,
// cards, the deck) lives in LDS, laid out [field][seat][lane] so that the 64
// lanes of a wave touch 64 consecutive words.  Only seat 0 looks the strategy
// up, so only its buckets are computed; the showdown walks the rank table for
// the seats still in.  Net chips of seat 0 are summed with one 64-bit atomic
// per wave into one of 64 counters (integer sums: the total does not depend
// on order).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/mccfr/mccfr.h"
#include "device_common.hpp"
#include "kernel_util.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace mccfr {
namespace {

constexpr int kEvalSeats = 9;           // eval.h:67 MAX_EVAL_PLAYERS
constexpr uint32_t kEvalTag = 0xE7A1u;  // Philox counter word 3 of the harness streams
constexpr int kEvalWG = 128;
constexpr uint64_t kEvalChunk = 1ull << 22;  // hands per launch

struct EvalArgs {
  const float* P;        // normalised strategy rows
  uint32_t slot_mask;
  int n;                 // seats
  int opp;               // archetype, 4 = random
  int stack;
  uint32_t seed_lo, seed_hi;
  uint64_t first_hand;
  uint32_t count;        // hands in this launch
  const int32_t* hr;
  const uint8_t* lut;
  const uint8_t* flop5;  // clamped harness buckets of every 5-card set
  const uint8_t* turn6;  // ... of every 6-card set
  int32_t* per_hand;     // [count] or null
  unsigned long long* sum;  // [kEvalSums][kEvalSumStride]: partial sums of seat 0's net chips
};
constexpr uint32_t kEvalSums = 64;
constexpr uint32_t kEvalSumStride = 16;

// {fold, call, raise} of the four scripted archetypes (eval.h:5-13).
__constant__ float c_arch[4][3] = {{0.08f, 0.72f, 0.20f}, {0.60f, 0.30f, 0.10f}, {0.10f, 0.20f, 0.70f},
                                   {0.28f, 0.44f, 0.28f}};

__device__ __forceinline__ float eval_uniform(DrawStream& s, float total) {
  return (float)(s.next() >> 8) * (1.0f / 16777216.0f) * total;
}

// eval.cpp:49-69 over a menu of A actions.
template <int ROW>
__device__ int sample_probs(const DevGame& g, const float (&probs)[ROW], uint32_t vm, DrawStream& rng) {
  float total = 0.f;
#pragma unroll
  for (int a = 0; a < ROW; ++a)
    if (vm >> a & 1u) total += probs[a];
  if (total < 1e-7f) return passive_fallback(g, vm);
  const float r = eval_uniform(rng, total);
  float cum = 0.f;
  int last = -1, pick = -1;
#pragma unroll
  for (int a = 0; a < ROW; ++a) {
    if (!(vm >> a & 1u)) continue;
    cum += probs[a];
    last = a;
    if (pick < 0 && r <= cum) pick = a;
  }
  return pick >= 0 ? pick : last;
}

// eval.cpp:100-157 by action kind: a free check takes the fold and call
// shares, raises (sized raises and all-in) split the raise share evenly.
template <int ROW>
__device__ int scripted(const DevGame& g, int opp, uint32_t vm, int to_call, DrawStream& rng) {
  if (opp == 4) {
    uint32_t k = rng.next() % (uint32_t)__popc(vm);
    uint32_t m = vm;
    for (; k; --k) m &= m - 1;
    return __builtin_ctz(m);
  }
  const float fold_p = c_arch[opp][0], call_p = c_arch[opp][1], raise_p = c_arch[opp][2];
  const uint32_t chk = vm & (to_call == 0 ? g.check_mask : 0u);
  const uint32_t cal = vm & (to_call != 0 ? g.call_mask : 0u);
  const uint32_t fld = vm & g.fold_mask;
  const uint32_t rse = vm & (g.raise_mask | g.allin_mask);
  const int check_a = chk ? __builtin_ctz(chk) : -1;
  const int call_a = cal ? __builtin_ctz(cal) : -1;
  const int fold_a = fld ? __builtin_ctz(fld) : -1;
  const int n_raise = __popc(rse);
  float probs[ROW];
#pragma unroll
  for (int a = 0; a < ROW; ++a) probs[a] = 0.f;
  const float per = n_raise > 0 ? raise_p / (float)n_raise : 0.f;
#pragma unroll
  for (int a = 0; a < ROW; ++a) {
    float p = 0.f;
    if (a == check_a) p = fold_p + call_p;
    if (check_a < 0 && a == fold_a) p = fold_p;
    if (check_a < 0 && a == call_a) p = call_p;
    if (n_raise > 0 && (rse >> a & 1u)) p = per;
    if (n_raise == 0) {
      const int gets = check_a >= 0 ? check_a : (call_a >= 0 ? call_a : fold_a);
      if (a == gets) p += raise_p;
    }
    probs[a] = p;
  }
  return sample_probs<ROW>(g, probs, vm, rng);
}

template <int ROW>
__global__ __launch_bounds__(kEvalWG) void k_eval_hands(const DevGame g, const EvalArgs e) {
  __shared__ int s_stack[kEvalSeats][kEvalWG];
  __shared__ int s_bet[kEvalSeats][kEvalWG];
  __shared__ int s_inv[kEvalSeats][kEvalWG];
  __shared__ uint8_t s_deck[52][kEvalWG];
  const int t = threadIdx.x;
  const uint32_t i = blockIdx.x * kEvalWG + t;
  int64_t net0 = 0;
  if (i < e.count) {
    const int n = e.n;
    const uint64_t hand = e.first_hand + i;
    const int dealer = (int)(hand % (uint64_t)n);
    DrawStream rng;
    rng.init(e.seed_lo, e.seed_hi, hand, kEvalTag);

    // deal: 2n+5 partial Fisher-Yates draws on a fresh deck
    for (int c = 0; c < 52; ++c) s_deck[c][t] = (uint8_t)c;
    for (int k = 0; k < 2 * n + 5; ++k) {
      const uint32_t r = rng.next();
      const int j = k + (int)(((uint64_t)r * (uint64_t)(52 - k)) >> 32);
      const uint8_t a = s_deck[k][t], b = s_deck[j][t];
      s_deck[k][t] = b;
      s_deck[j][t] = a;
    }
    const uint32_t b0 = s_deck[2 * n][t], b1 = s_deck[2 * n + 1][t], b2 = s_deck[2 * n + 2][t],
                   b3 = s_deck[2 * n + 3][t], b4 = s_deck[2 * n + 4][t];

    // seat 0's buckets (abstraction.cpp:191-207)
    uint32_t bk[4];
    const uint32_t h0 = s_deck[0][t], h1 = s_deck[1][t];
    const uint32_t hb0 = e.lut[h0 * 52 + h1];
    bk[0] = 0;
    bk[1] = e.flop5[colex5(h0, h1, b0, b1, b2)];
    bk[2] = e.turn6[colex6(h0, h1, b0, b1, b2, b3)];
    {
      int p = e.hr[53 + card_rm(h0)];
      p = e.hr[p + card_rm(h1)];
      p = e.hr[p + card_rm(b0)];
      p = e.hr[p + card_rm(b1)];
      p = e.hr[p + card_rm(b2)];
      p = e.hr[p + card_rm(b3)];
      p = e.hr[p + card_rm(b4)];
      bk[3] = eval_bucket_of_rank(decode_packed((uint32_t)p & 0xFFFFu));
    }

    // blinds (eval.cpp:499-517)
    for (int p = 0; p < n; ++p) {
      s_stack[p][t] = e.stack;
      s_bet[p][t] = 0;
      s_inv[p][t] = 0;
    }
    uint32_t folded = 0, allin = 0;
    int pot = 0;
    const int sb_pos = n == 2 ? dealer : (dealer + 1 < n ? dealer + 1 : dealer + 1 - n);
    const int bb_pos = n == 2 ? (dealer + 1 < n ? dealer + 1 : 0) : (dealer + 2) % n;
    {
      const int c0 = e.stack < g.sb ? e.stack : g.sb;
      s_stack[sb_pos][t] -= c0;
      s_bet[sb_pos][t] = c0;
      s_inv[sb_pos][t] = c0;
      pot += c0;
      if (!s_stack[sb_pos][t]) allin |= 1u << sb_pos;
      const int c1 = e.stack < g.bb ? e.stack : g.bb;
      s_stack[bb_pos][t] -= c1;
      s_bet[bb_pos][t] = c1;
      s_inv[bb_pos][t] = c1;
      pot += c1;
      if (!s_stack[bb_pos][t]) allin |= 1u << bb_pos;
    }
    int current_bet = s_bet[bb_pos][t];
    int lfr = g.bb, raises = 0;
    uint32_t ab = 0;
    const uint32_t all = (1u << n) - 1u;

    // streets (eval.cpp:535-560)
    for (int st = 0; st < 4; ++st) {
      if (st > 0) {
        for (int p = 0; p < n; ++p) s_bet[p][t] = 0;
        current_bet = 0;
        lfr = g.bb;
        raises = 0;
      }
      const uint32_t live = all & ~folded;
      if (__popc(live) <= 1) break;
      if (__popc(live & ~allin) <= 1) {
        if (st == 3) break;
        continue;
      }
      int first = st == 0 ? (bb_pos + 1) % n : (dealer + 1) % n;
      if (st > 0)
        for (int k = 0; k < n; ++k) {
          const int q = (dealer + 1 + k) % n;
          if ((live & ~allin) >> q & 1u) { first = q; break; }
        }

      // betting round (eval.cpp:322-391)
      uint32_t need = live & ~allin;
      int p = first;
      bool hand_over = false;
      const int safety = n * n * 4;
      for (int guard = 0; guard < safety && need; ++guard) {
        if (!(need >> p & 1u)) {
          p = p + 1 == n ? 0 : p + 1;
          continue;
        }
        const int stack = s_stack[p][t], bet = s_bet[p][t];
        const int to_call = current_bet - bet;
        const uint32_t vm = valid_mask(g, pot, stack, to_call, current_bet, lfr, raises);
        int chosen;
        if (p == 0) {
          // eval.cpp:85-97: the strategy row, if get_strategy would keep it
          const uint32_t key = info_hash(0u, hb0, bk[st], (uint32_t)st, ab, e.slot_mask);
          float probs[ROW];
          load_f32_row<ROW>(e.P + (size_t)key * ROW, probs);
          float tot = 0.f;
#pragma unroll
          for (int a = 0; a < ROW; ++a) tot += probs[a];
          chosen = tot >= 1e-6f ? sample_probs<ROW>(g, probs, vm, rng) : passive_fallback(g, vm);
        } else {
          chosen = scripted<ROW>(g, e.opp, vm, to_call, rng);
        }
        ab = ((ab << g.hist_bits) | (uint32_t)chosen) & 0x3FFFFFFFu;
        need &= ~(1u << p);
        const int kind = g.kind[chosen];
        if (kind == MCCFR_ACT_FOLD) {
          folded |= 1u << p;
          if (__popc(all & ~folded) <= 1) { hand_over = true; break; }
        } else if (kind != MCCFR_ACT_CHECK) {
          const int chips = chips_for(g, chosen, pot, stack, to_call, current_bet, lfr);
          const int prev = current_bet;
          s_stack[p][t] = stack - chips;
          s_bet[p][t] = bet + chips;
          s_inv[p][t] += chips;
          pot += chips;
          if (stack - chips == 0) allin |= 1u << p;
          if (bet + chips > current_bet) {
            const int raise = bet + chips - prev;
            if (raise >= lfr || prev == 0) lfr = raise;
            current_bet = bet + chips;
            ++raises;
            need = all & ~folded & ~allin & ~(1u << p);
          }
        }
        p = p + 1 == n ? 0 : p + 1;
      }
      if (hand_over) break;
      if (__popc(all & ~folded) <= 1 || st == 3) break;
    }

    // payoff of seat 0 (eval.cpp:563-571, 410-474)
    const uint32_t live = all & ~folded;
    if (__popc(live) == 1) {
      net0 = (live & 1u ? pot : 0) - s_inv[0][t];
    } else {
      // ranks of the live seats; s_bet holds the investment still to settle
      uint32_t rank[kEvalSeats];
      for (int q = 0; q < kEvalSeats; ++q) rank[q] = 0;
      for (int q = 0; q < n; ++q) {
        s_bet[q][t] = s_inv[q][t];
        if (!(live >> q & 1u)) continue;
        int w = e.hr[53 + card_rm(s_deck[2 * q][t])];
        w = e.hr[w + card_rm(s_deck[2 * q + 1][t])];
        w = e.hr[w + card_rm(b0)];
        w = e.hr[w + card_rm(b1)];
        w = e.hr[w + card_rm(b2)];
        w = e.hr[w + card_rm(b3)];
        w = e.hr[w + card_rm(b4)];
        const uint32_t r = decode_packed((uint32_t)w & 0xFFFFu);
#pragma unroll
        for (int z = 0; z < kEvalSeats; ++z)
          if (z == q) rank[z] = r;
      }
      int won0 = 0;
      while (true) {
        int lvl = 0x7FFFFFFF;
        for (int q = 0; q < n; ++q) {
          const int rem = s_bet[q][t];
          if (rem > 0 && rem < lvl) lvl = rem;
        }
        if (lvl == 0x7FFFFFFF) break;
        int side = 0;
        uint32_t elig = 0;
        for (int q = 0; q < n; ++q) {
          const int rem = s_bet[q][t];
          if (rem <= 0) continue;
          const int c = rem < lvl ? rem : lvl;
          side += c;
          s_bet[q][t] = rem - c;
          if (live >> q & 1u) elig |= 1u << q;
        }
        uint32_t best = 0;
#pragma unroll
        for (int q = 0; q < kEvalSeats; ++q)
          if ((elig >> q & 1u) && rank[q] > best) best = rank[q];
        int nw = 0;
#pragma unroll
        for (int q = 0; q < kEvalSeats; ++q) nw += (elig >> q & 1u) && rank[q] == best;
        if (nw == 0) continue;
        if ((elig & 1u) && rank[0] == best) won0 += side / nw;
      }
      net0 = won0 - s_inv[0][t];
    }
    if (e.per_hand) e.per_hand[i] = (int32_t)net0;
  }
  // one 64-bit atomic per wave, on one of kEvalSums counters (128 bytes apart:
  // atomics on one address serialize across the device); the host adds them
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) net0 += __shfl_xor(net0, off, 64);
  if ((t & 63) == 0 && net0 != 0)
    atomicAdd(e.sum + (size_t)((t >> 6) % kEvalSums) * kEvalSumStride, (unsigned long long)net0);
}

__global__ __launch_bounds__(256) void k_scatter_rows(const uint32_t* __restrict__ keys,
                                                      const float* __restrict__ probs, uint32_t n_rows, int row,
                                                      uint64_t slots, float* __restrict__ P) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n_rows; r += gridDim.x * blockDim.x) {
    const uint64_t k = keys[r];
    if (k >= slots) continue;
    for (int a = 0; a < row; ++a) P[k * row + a] = probs[(size_t)r * row + a];
  }
}

int ensure_eval_buckets(mccfr_ctx* c) {
  if (c->eval_bk_ready) return MCCFR_OK;
  if (!c->eval_bk.flop5) HIPX(c, hipMalloc(&c->eval_bk.flop5, kFlopSets));
  if (!c->eval_bk.turn6) HIPX(c, hipMalloc(&c->eval_bk.turn6, kTurnSets));
  launch_build_buckets(c->stream, c->hr, c->eval_bk, true);
  HIPX(c, hipGetLastError());
  c->eval_bk_ready = true;
  return MCCFR_OK;
}

// Plays the hands on the normalised table P (device, slots x row).
int run_eval(mccfr_ctx* c, const DevGame& g, const float* P, int n_players, int opponent, int64_t n_hands,
             uint64_t seed, int32_t* per_hand, mccfr_eval_result* out) {
  int rc = ensure_eval_buckets(c);
  if (rc) return rc;
  unsigned long long* d_sum = nullptr;
  int32_t* d_hand = nullptr;
  constexpr size_t kSumBytes = (size_t)kEvalSums * kEvalSumStride * sizeof(unsigned long long);
  HIPX(c, hipMalloc(&d_sum, kSumBytes));
  hipError_t err = per_hand && n_hands > 0
                       ? hipMalloc(&d_hand, (size_t)std::min<uint64_t>(kEvalChunk, (uint64_t)n_hands) * 4)
                       : hipSuccess;
  if (err == hipSuccess) err = hipMemsetAsync(d_sum, 0, kSumBytes, c->stream);
  EvalArgs e{};
  e.P = P;
  e.slot_mask = (uint32_t)((1ull << g.table_log2) - 1ull);
  e.n = n_players;
  e.opp = opponent;
  e.stack = g.stack[0];
  e.seed_lo = (uint32_t)seed;
  e.seed_hi = (uint32_t)(seed >> 32);
  e.hr = c->hr;
  e.lut = c->lut;
  e.flop5 = c->eval_bk.flop5;
  e.turn6 = c->eval_bk.turn6;
  e.sum = d_sum;
  for (uint64_t done = 0; err == hipSuccess && done < (uint64_t)n_hands; done += kEvalChunk) {
    const uint32_t cnt = (uint32_t)std::min<uint64_t>(kEvalChunk, (uint64_t)n_hands - done);
    e.first_hand = done;
    e.count = cnt;
    e.per_hand = d_hand;
    const dim3 grid((cnt + kEvalWG - 1) / kEvalWG);
    if (g.row == 8)
      hipLaunchKernelGGL(k_eval_hands<8>, grid, dim3(kEvalWG), 0, c->stream, g, e);
    else
      hipLaunchKernelGGL(k_eval_hands<16>, grid, dim3(kEvalWG), 0, c->stream, g, e);
    err = hipGetLastError();
    if (err == hipSuccess && per_hand)
      err = hipMemcpyAsync(per_hand + done, d_hand, (size_t)cnt * 4, hipMemcpyDeviceToHost, c->stream);
  }
  unsigned long long parts[kEvalSums * kEvalSumStride] = {};
  if (err == hipSuccess) err = hipMemcpyAsync(parts, d_sum, kSumBytes, hipMemcpyDeviceToHost, c->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
  unsigned long long sum = 0;  // (mod 2^64: the two's-complement total)
  for (uint32_t k = 0; k < kEvalSums; ++k) sum += parts[k * kEvalSumStride];
  (void)hipFree(d_sum);
  if (d_hand) (void)hipFree(d_hand);
  if (err != hipSuccess) return fail(c, MCCFR_E_HIP, hipGetErrorString(err));
  out->hands_played = n_hands;
  out->net_chips = (int64_t)sum;
  out->net_bb = (double)out->net_chips / (double)g.bb;
  out->bb_per_100 = n_hands ? out->net_bb / (double)n_hands * 100.0 : 0.0;
  return MCCFR_OK;
}

int check_eval_args(mccfr_ctx* c, int n_players, int opponent, int64_t n_hands, const mccfr_eval_result* out) {
  if (!c || !out) return fail(c, MCCFR_E_ARG, "null argument");
  if (n_players < 2 || n_players > kEvalSeats) return fail(c, MCCFR_E_ARG, "n_players must be 2..9");
  if (opponent < 0 || opponent > 4) return fail(c, MCCFR_E_ARG, "opponent must be 0..4");
  if (n_hands < 0) return fail(c, MCCFR_E_ARG, "n_hands must be >= 0");
  if (!c->hr) return fail(c, MCCFR_E_STATE, "no hand table loaded");
  return MCCFR_OK;
}

}  // namespace
}  // namespace mccfr

using namespace mccfr;

extern "C" {

int mccfr_evaluate_strategy(mccfr_ctx* c, uint32_t tenant, int n_players, int opponent, int64_t n_hands,
                            uint64_t seed, int32_t* per_hand, mccfr_eval_result* out) {
  int rc = check_eval_args(c, n_players, opponent, n_hands, out);
  if (rc) return rc;
  if (tenant >= c->tenants.size() || !c->tenants[tenant].used) return fail(c, MCCFR_E_ARG, "no such tenant");
  const Tenant& t = c->tenants[tenant];
  if (t.hg.root_street != 0) return fail(c, MCCFR_E_ARG, "the harness plays whole hands: tenant has a postflop root");
  float* P = nullptr;
  HIPX(c, hipMalloc(&P, t.elems() * sizeof(float)));
  launch_normalize(c->stream, t.row, t.hg.S, P, t.slots);
  DevGame g = t.hg;
  g.stack[0] = t.cfg.starting_stack;
  rc = run_eval(c, g, P, n_players, opponent, n_hands, seed, per_hand, out);
  (void)hipFree(P);
  return rc;
}

int mccfr_evaluate_rows(mccfr_ctx* c, const mccfr_game_cfg* cfg, const uint32_t* keys, const float* probs,
                        uint32_t n_rows, uint32_t num_actions, int n_players, int opponent, int64_t n_hands,
                        uint64_t seed, int32_t* per_hand, mccfr_eval_result* out) {
  int rc = check_eval_args(c, n_players, opponent, n_hands, out);
  if (rc) return rc;
  if (!cfg || (n_rows && (!keys || !probs))) return fail(c, MCCFR_E_ARG, "null argument");
  DevGame g{};
  const char* why = nullptr;
  if (!make_dev_game(*cfg, nullptr, g, &why)) return fail(c, MCCFR_E_ARG, why);
  if ((int)num_actions != g.row)
    return fail(c, MCCFR_E_ARG, "rows must hold " + std::to_string(g.row) + " probabilities for this game");
  const uint64_t slots = 1ull << g.table_log2;
  const size_t elems = (size_t)slots * g.row;
  float* P = nullptr;
  uint32_t* dk = nullptr;
  float* dp = nullptr;
  HIPX(c, hipMalloc(&P, elems * sizeof(float)));
  hipError_t err = hipSuccess;
  if (n_rows) {
    err = hipMalloc(&dk, (size_t)n_rows * 4);
    if (err == hipSuccess) err = hipMalloc(&dp, (size_t)n_rows * g.row * 4);
    if (err == hipSuccess) err = hipMemcpyAsync(dk, keys, (size_t)n_rows * 4, hipMemcpyHostToDevice, c->stream);
    if (err == hipSuccess)
      err = hipMemcpyAsync(dp, probs, (size_t)n_rows * g.row * 4, hipMemcpyHostToDevice, c->stream);
  }
  if (err == hipSuccess) {
    launch_zero(c->stream, P, elems * sizeof(float));
    if (n_rows)
      hipLaunchKernelGGL(k_scatter_rows, dim3(grid_for(n_rows, 256, 65536)), dim3(256), 0, c->stream, dk, dp,
                         n_rows, g.row, slots, P);
    err = hipGetLastError();
  }
  if (err == hipSuccess) {
    g.stack[0] = cfg->starting_stack;
    rc = run_eval(c, g, P, n_players, opponent, n_hands, seed, per_hand, out);
  } else {
    rc = fail(c, MCCFR_E_HIP, hipGetErrorString(err));
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(P);
  if (dk) (void)hipFree(dk);
  if (dp) (void)hipFree(dp);
  return rc;
}

}  // extern "C"
