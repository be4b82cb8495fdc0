// oracle.cpp — CPU restatement of the reference MCCFR hot path.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker / timed baseline; never by the
// product library.  Built with -ffp-contract=off so that every FP32 operation
// is rounded exactly where the reference's source writes it.
//
// Parity pins (each function is checked against something other than itself):
//   orc_walk5/6/7, orc_decode_packed, orc_validate_table
//       the reference's own hand_eval.cpp compiled under oracle/_ref (its loader
//       validates the generated table and its evaluate_7cards/evaluate_best
//       agree word for word); royal-flush KAT (hand_eval.cpp:70-84);
//       exhaustive 7-card category counts (textbook); brute force (below).
//   orc_bf_class*, orc_exhaustive5_bf
//       textbook 5-card class counts per category (1277/2860/858/858/10/1277/
//       156/156/10) and 5-card hand counts (1,302,540 ... 40).
//   orc_philox4x32_10     Random123 known-answer vectors; rocRAND's Philox.
//   orc_f32_to_bf16       IEEE round-to-nearest-even, checked against numpy
//                         float32 -> bfloat16 rounding done in integer math.
//   orc_info_hash         FNV-1a spec (offset 2166136261, prime 16777619).
//   orc_preflop_lut, orc_valid_mask, orc_chips
//       golden fixtures written by the reference's abstraction.cpp (oracle/_ref).
//   orc_deal              permutation invariants (distinct cards) + uniformity.
//   orc_replay_launch, orc_walk_cards (the Walker: rules, apply, queue rebuild,
//   raise cap, history, side pots, depth guard, update-node contributions)
//       nine hand-worked forced-play sequences (tests/forced_play.py: every
//       decision's player, history, legal actions, chips, pot, bet to match,
//       last full raise, raise count and queue, the terminal utilities and the
//       contributions written as literals; one-hot strategy rows force the
//       play), covering open / re-raise / call queue order, a short all-in
//       that keeps the last full raise, min-raises, the raise cap on the 4- and
//       10-action menus, 3- and 4-bit history and its 30-bit mask, a three-way
//       all-in with side pots, heads-up order, fold-around and the depth guard;
//       brute force of the 2-player games with stacks <= the big blind (all-in
//       from the blinds; {fold, all-in} for the small blind) against the
//       brute-force evaluator; the first-launch closed form (all-zero tables:
//       passive fallback everywhere, no draws); chip conservation; the
//       sum_a sigma(a)*rdelta(a) = 0 invariant.
//   orc_replay_launch with sampling = external (every legal action at the
//   update player's nodes, values returned up the tree)
//       brute force of the heads-up 60-chip game on the 8-action menu: the
//       whole tree enumerated, exact counterfactual values and expected
//       contributions; equal per traversal against one-hot opponents, equal
//       in the mean over many traversals against mixed ones; the
//       sum_a sigma(a)*rdelta(a) = 0 invariant per update node.
//   orc_cpu_baseline      same contributions as orc_replay_launch (test).
// Parity unpinned: cuRAND streams and __float2bfloat16 (replaced by the
// project's own Philox and RNE definitions), and eval5/eval6 against the
// reference's downloaded HandRanks.dat (absent here).
#include "oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

namespace {

// ===========================================================================
// Cards and table walks.
// ===========================================================================
inline int to_rm(int c) { return (c % 13) * 4 + (c / 13) + 1; }  // hand_eval.cpp:23-27

inline uint16_t decode_packed(uint16_t packed) {  // hand_eval.cpp:29-45
  const uint16_t cat = packed >> 12;
  const uint16_t ord = packed & 0x0FFFu;
  if (ord == 0) return 0;
  static const uint16_t off[10] = {0, 0, 1277, 4137, 4995, 5853, 5863, 7140, 7296, 7452};
  if (cat < 1 || cat > 9) return 0;
  return (uint16_t)(off[cat] + ord);
}

// hand_eval.cpp:107-113 / cfr_gpu.cu:114-122: five lookups, decode of the last word read.
inline uint16_t walk5(const int32_t* hr, const uint8_t* c) {
  int p = hr[53 + to_rm(c[0])];
  p = hr[p + to_rm(c[1])];
  p = hr[p + to_rm(c[2])];
  p = hr[p + to_rm(c[3])];
  return decode_packed((uint16_t)hr[p + to_rm(c[4])]);
}
// hand_eval.cpp:119-127 / cfr_gpu.cu:140-150
inline uint16_t walk6(const int32_t* hr, const uint8_t* c) {
  int p = hr[53 + to_rm(c[0])];
  for (int i = 1; i < 5; ++i) p = hr[p + to_rm(c[i])];
  return decode_packed((uint16_t)hr[p + to_rm(c[5])]);
}
// hand_eval.cpp:134-143 / cfr_gpu.cu:126-137
inline uint16_t walk7(const int32_t* hr, const uint8_t* c) {
  int p = hr[53 + to_rm(c[0])];
  for (int i = 1; i < 6; ++i) p = hr[p + to_rm(c[i])];
  return decode_packed((uint16_t)hr[p + to_rm(c[6])]);
}

inline int category_of_class(int cls) {  // hand_eval.cpp:194-215 boundaries
  static const int hi[9] = {1277, 4137, 4995, 5853, 5863, 7140, 7296, 7452, 7462};
  for (int c = 0; c < 9; ++c)
    if (cls <= hi[c]) return c + 1;
  return 0;
}

// ===========================================================================
// Brute-force evaluator: classify each 5-card subset by sorting, order the
// classes by enumerating all 2,598,960 five-card hands.
// ===========================================================================
constexpr uint32_t P13[6] = {1, 13, 169, 2197, 28561, 371293};

uint32_t bf_key5(const int* r, const int* s) {
  int cnt[13] = {0};
  for (int i = 0; i < 5; ++i) ++cnt[r[i]];
  // groups ordered by (count desc, rank desc)
  int gc[5], gr[5], ng = 0;
  for (int want = 4; want >= 1; --want)
    for (int rk = 12; rk >= 0; --rk)
      if (cnt[rk] == want) { gc[ng] = want; gr[ng] = rk; ++ng; }
  const bool flush = s[0] == s[1] && s[1] == s[2] && s[2] == s[3] && s[3] == s[4];
  bool straight = false;
  int top = 0;
  if (ng == 5) {
    if (gr[0] - gr[4] == 4) { straight = true; top = gr[0]; }
    else if (gr[0] == 12 && gr[1] == 3 && gr[2] == 2 && gr[3] == 1 && gr[4] == 0) { straight = true; top = 3; }
  }
  int cat;
  if (straight && flush) cat = 8;
  else if (gc[0] == 4) cat = 7;
  else if (gc[0] == 3 && gc[1] == 2) cat = 6;
  else if (flush) cat = 5;
  else if (straight) cat = 4;
  else if (gc[0] == 3) cat = 3;
  else if (gc[0] == 2 && gc[1] == 2) cat = 2;
  else if (gc[0] == 2) cat = 1;
  else cat = 0;
  uint32_t key = (uint32_t)cat * P13[5];
  if (straight) return key + (uint32_t)top * P13[4];
  for (int i = 0; i < ng; ++i) key += (uint32_t)gr[i] * P13[4 - i];
  return key;
}

struct BFClasses {
  std::vector<uint16_t> cls;  // key -> class 1..7462
  int count = 0;
  int per_cat[9] = {0};
  BFClasses() : cls(9 * P13[5], 0) {
    std::vector<uint8_t> seen(9 * P13[5], 0);
    int r[5], s[5];
    for (int a = 0; a < 52; ++a)
      for (int b = a + 1; b < 52; ++b)
        for (int c = b + 1; c < 52; ++c)
          for (int d = c + 1; d < 52; ++d)
            for (int e = d + 1; e < 52; ++e) {
              const int cs[5] = {a, b, c, d, e};
              for (int i = 0; i < 5; ++i) { r[i] = cs[i] % 13; s[i] = cs[i] / 13; }
              seen[bf_key5(r, s)] = 1;
            }
    for (uint32_t k = 0; k < seen.size(); ++k)
      if (seen[k]) {
        cls[k] = (uint16_t)++count;
        ++per_cat[k / P13[5]];
      }
  }
};

const BFClasses& bf() {
  static const BFClasses b;
  return b;
}

int bf_class_n(const uint8_t* cards, int n) {
  static const int idx7[21][5] = {
      {0, 1, 2, 3, 4}, {0, 1, 2, 3, 5}, {0, 1, 2, 3, 6}, {0, 1, 2, 4, 5}, {0, 1, 2, 4, 6},
      {0, 1, 2, 5, 6}, {0, 1, 3, 4, 5}, {0, 1, 3, 4, 6}, {0, 1, 3, 5, 6}, {0, 1, 4, 5, 6},
      {0, 2, 3, 4, 5}, {0, 2, 3, 4, 6}, {0, 2, 3, 5, 6}, {0, 2, 4, 5, 6}, {0, 3, 4, 5, 6},
      {1, 2, 3, 4, 5}, {1, 2, 3, 4, 6}, {1, 2, 3, 5, 6}, {1, 2, 4, 5, 6}, {1, 3, 4, 5, 6},
      {2, 3, 4, 5, 6}};
  const BFClasses& B = bf();
  int best = 0;
  int r[5], s[5];
  for (int k = 0; k < 21; ++k) {
    bool ok = true;
    for (int i = 0; i < 5; ++i) {
      const int j = idx7[k][i];
      if (j >= n) { ok = false; break; }
      r[i] = cards[j] % 13;
      s[i] = cards[j] / 13;
    }
    if (!ok) continue;
    best = std::max<int>(best, B.cls[bf_key5(r, s)]);
  }
  return best;
}

// ===========================================================================
// Philox4x32-10 and the draw / rounding conventions of the project.
// ===========================================================================
inline void philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Draw k of traversal `tid`: word k&3 of Philox(counter = {k>>2, tid_lo, tid_hi, tag},
// key = {seed_lo, seed_hi}).
// `path` (external sampling) names the task of the traversal the stream
// belongs to: 0 for the root (and every outcome-sampling walk), the task's
// path code for the subtree under an update-player action; it occupies the
// counter's first word above the block index (k < 256 on every walk).
struct Stream {
  uint32_t key[2];
  uint32_t tid_lo, tid_hi, tag;
  uint32_t path = 0;
  uint32_t k = 0;
  uint32_t blk[4];
  uint32_t blk_id = 0xFFFFFFFFu;
  Stream(uint64_t seed, uint64_t tid, uint32_t tg)
      : key{(uint32_t)seed, (uint32_t)(seed >> 32)}, tid_lo((uint32_t)tid),
        tid_hi((uint32_t)(tid >> 32)), tag(tg) {}
  uint32_t next() {
    const uint32_t b = k >> 2;
    if (b != blk_id) {
      const uint32_t ctr[4] = {b | (path << 6), tid_lo, tid_hi, tag};
      philox(ctr, key, blk);
      blk_id = b;
    }
    return blk[k++ & 3];
  }
};

// Path code of the task under action a of an update-player node whose own
// task has code p (external sampling): FNV-1a step, 26 bits, never 0.
inline uint32_t es_child_path(uint32_t p, int a) {
  const uint32_t h = ((p ^ (uint32_t)(a + 1)) * 16777619u) & 0x03FFFFFFu;
  return h ? h : 1u;
}

// FNV-1a of one (slot, action) pair from the offset basis: an external-
// sampling trace sums these (mod 2^32) over every entered (node, action), an
// order-free check of the explored tree.
inline uint32_t fnv_pair(uint32_t slot, int a) {
  uint32_t h = 2166136261u;
  for (int i = 0; i < 4; ++i) h = (h ^ ((slot >> (8 * i)) & 0xFFu)) * 16777619u;
  return (h ^ (uint32_t)(a & 0xFF)) * 16777619u;
}

// Uniform in (0, 1]: ((x >> 8) + 1) * 2^-24, exact in FP32.
inline float u01(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

inline uint16_t f2bf(float x) {  // IEEE binary32 -> bfloat16, round to nearest even
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
inline float bf2f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// ===========================================================================
// Abstraction.
// ===========================================================================
// abstraction.cpp:10-37 (Chen-formula variant; rank 0 = deuce .. 12 = ace)
float chen_score(int h0, int h1) {
  int r0 = h0 % 13, r1 = h1 % 13;
  const bool suited = (h0 / 13) == (h1 / 13);
  if (r0 < r1) std::swap(r0, r1);
  float base;
  if (r0 == 12) base = 10.f;
  else if (r0 == 11) base = 8.f;
  else if (r0 == 10) base = 7.f;
  else if (r0 == 9) base = 6.f;
  else base = std::max(0.5f, r0 / 2.f);
  float score;
  if (r0 == r1) {
    score = std::max(5.f, base * 2.f);
  } else {
    score = base;
    if (suited) score += 2.f;
    const int gap = r0 - r1 - 1;
    if (gap == 1) score -= 1.f;
    else if (gap == 2) score -= 2.f;
    else if (gap == 3) score -= 4.f;
    else if (gap >= 4) score -= 5.f;
    if (gap <= 1 && r0 < 12) score += 1.f;
  }
  return score;
}

// abstraction.cpp:42-59 + cfr_gpu.cu:1164-1173 (diagonal = 0)
void preflop_lut(uint8_t* lut) {
  std::vector<float> sorted;
  for (int i = 0; i < 52; ++i)
    for (int j = i + 1; j < 52; ++j) sorted.push_back(chen_score(i, j));
  std::sort(sorted.begin(), sorted.end());
  for (int h0 = 0; h0 < 52; ++h0)
    for (int h1 = 0; h1 < 52; ++h1) {
      if (h0 == h1) { lut[h0 * 52 + h1] = 0; continue; }
      const float s = chen_score(h0, h1);
      const size_t idx = (size_t)(std::lower_bound(sorted.begin(), sorted.end(), s) - sorted.begin());
      const float pct = (float)idx / sorted.size();
      lut[h0 * 52 + h1] = (uint8_t)std::min((int)(pct * 64), 63);
    }
}

// cfr_gpu.cu:165-179 with the slot mask generalised to table_log2 bits.
inline uint32_t info_hash(uint8_t player, uint8_t hb, uint8_t bb, uint8_t street, uint32_t ab,
                          int log2) {
  uint32_t h = 2166136261u;
  h = (h ^ player) * 16777619u;
  h = (h ^ hb) * 16777619u;
  h = (h ^ bb) * 16777619u;
  h = (h ^ street) * 16777619u;
  h = (h ^ (uint8_t)(ab & 0xFF)) * 16777619u;
  h = (h ^ (uint8_t)((ab >> 8) & 0xFF)) * 16777619u;
  h = (h ^ (uint8_t)((ab >> 16) & 0xFF)) * 16777619u;
  h = (h ^ (uint8_t)((ab >> 24) & 0xFF)) * 16777619u;
  return h & (uint32_t)((1ull << log2) - 1ull);
}

inline int min_raise_to(int current_bet, int lfr, int bb) {  // cfr_gpu.cu:203-205
  return current_bet == 0 ? std::max(1, bb) : current_bet + std::max(lfr, bb);
}

inline int frac_target(const mccfr_action& a, int pot, int current_bet, int mrt) {
  return std::max(mrt, current_bet + std::max(1, (a.num * pot) / a.den));
}

// cfr_gpu.cu:186-215 over a general menu; a street's raise cap (0 = none)
// removes every raise and the raising all-in once `raises` reaches it.
uint32_t valid_mask(const mccfr_game_cfg& g, int pot, int stack, int to_call, int current_bet,
                    int lfr, int raises) {
  if (pot <= 0) pot = 1;
  uint32_t mask = 0;
  const int A = g.num_actions;
  auto set_kind = [&](int kind) {
    for (int a = 0; a < A; ++a)
      if (g.menu[a].kind == kind) mask |= 1u << a;
  };
  if (to_call == 0) {
    set_kind(MCCFR_ACT_CHECK);
    set_kind(MCCFR_ACT_CHECK_CALL);
  } else if (to_call >= stack) {
    set_kind(MCCFR_ACT_FOLD);
    set_kind(MCCFR_ACT_ALL_IN);
    return mask;
  } else {
    set_kind(MCCFR_ACT_FOLD);
    set_kind(MCCFR_ACT_CALL);
    set_kind(MCCFR_ACT_CHECK_CALL);
  }
  const int head = stack - to_call;
  if (head <= 0) return mask;
  const bool capped = g.max_raises > 0 && raises >= g.max_raises;
  if (capped) return mask;
  const int mrt = min_raise_to(current_bet, lfr, g.big_blind);
  const int maxrt = current_bet + head;
  if (maxrt >= mrt) {
    for (int a = 0; a < A; ++a) {
      const mccfr_action& m = g.menu[a];
      if (m.kind == MCCFR_ACT_RAISE_MIN) {
        if (mrt <= maxrt) mask |= 1u << a;
      } else if (m.kind == MCCFR_ACT_POT_FRAC) {
        if (frac_target(m, pot, current_bet, mrt) <= maxrt) mask |= 1u << a;
      }
    }
  }
  set_kind(MCCFR_ACT_ALL_IN);
  return mask;
}

// cfr_gpu.cu:217-244 over a general menu
int chips_for(const mccfr_game_cfg& g, int a, int pot, int stack, int to_call, int current_bet,
              int lfr) {
  if (pot <= 0) pot = 1;
  const int mrt = min_raise_to(current_bet, lfr, g.big_blind);
  const int maxrt = current_bet + std::max(0, stack - to_call);
  const mccfr_action& m = g.menu[a];
  switch (m.kind) {
    case MCCFR_ACT_FOLD:
    case MCCFR_ACT_CHECK: return 0;
    case MCCFR_ACT_CALL: return std::min(to_call, stack);
    case MCCFR_ACT_CHECK_CALL: return to_call == 0 ? 0 : std::min(to_call, stack);
    case MCCFR_ACT_RAISE_MIN:
    case MCCFR_ACT_POT_FRAC: {
      int target = m.kind == MCCFR_ACT_RAISE_MIN ? mrt : frac_target(m, pot, current_bet, mrt);
      target = std::min(target, maxrt);
      return std::min(to_call + std::max(0, target - current_bet), stack);
    }
    case MCCFR_ACT_ALL_IN: return stack;
    default: return 0;
  }
}

// cfr_gpu.cu:498-508: check > call > fold > raises in menu order > all-in.
int passive_fallback(const mccfr_game_cfg& g, uint32_t mask) {
  const int A = g.num_actions;
  static const int order[4] = {MCCFR_ACT_CHECK, MCCFR_ACT_CALL, MCCFR_ACT_CHECK_CALL, MCCFR_ACT_FOLD};
  for (int k : order)
    for (int a = 0; a < A; ++a)
      if ((mask >> a & 1u) && g.menu[a].kind == k) return a;
  for (int a = 0; a < A; ++a)
    if ((mask >> a & 1u) &&
        (g.menu[a].kind == MCCFR_ACT_RAISE_MIN || g.menu[a].kind == MCCFR_ACT_POT_FRAC))
      return a;
  for (int a = 0; a < A; ++a)
    if ((mask >> a & 1u) && g.menu[a].kind == MCCFR_ACT_ALL_IN) return a;
  return 0;
}

int row_width(const mccfr_game_cfg& g) { return g.num_actions <= 8 ? 8 : 16; }

// cfr_gpu.cu:476-489 (uint8 truncation of the bucket, not abstraction.cpp's clamp)
inline uint8_t bucket_of(uint16_t rank) { return (uint8_t)((uint32_t)(rank - 1) * 128u / 7462u); }

// ===========================================================================
// Traversal (cfr_gpu.cu:375-454, 498-860, 878-990), one lane at a time.
// ===========================================================================
struct Game {
  int N;
  uint8_t hole[12], comm[5], hb[6], bbk[6][4];
  int stacks[6], bets[6], invested[6];
  uint32_t folded, all_in;
  int pot, current_bet, lfr, bb_amt;
  uint32_t ab;
  int street;
  int to_act[6], n_to_act;
  int dealer;
  int raises;
};

// Reads of the pre-launch tables; both layouts restate a24 (cfr_gpu.cuh:90-95).
struct TablesAoS {
  const float* R;
  const uint16_t* sig;
  int row;
  float r(uint32_t slot, int a) const { return R[(size_t)slot * row + a]; }
  float s(uint32_t slot, int a) const { return bf2f(sig[(size_t)slot * row + a]); }
};
struct TablesSoA {
  const float* R;
  const uint16_t* sig;
  uint64_t slots;
  float r(uint32_t slot, int a) const { return R[(size_t)a * slots + slot]; }
  float s(uint32_t slot, int a) const { return bf2f(sig[(size_t)a * slots + slot]); }
};

struct SinkF64 {
  double *dR, *dS, *aR, *aS;
  int row;
  void add(uint32_t slot, int a, float rd, float sd) {
    const size_t i = (size_t)slot * row + a;
    dR[i] += (double)rd;
    dS[i] += (double)sd;
    aR[i] += std::fabs((double)rd);
    aS[i] += std::fabs((double)sd);
  }
};
struct SinkNull {
  void add(uint32_t, int, float, float) {}
};
struct Contribution {
  uint32_t slot;
  uint32_t a;
  float rd, sd;
};
struct SinkList {
  std::vector<Contribution>* out;
  void add(uint32_t slot, int a, float rd, float sd) { out->push_back({slot, (uint32_t)a, rd, sd}); }
};

template <class Tables, class Sink>
struct Walker {
  const mccfr_game_cfg& cfg;
  const int32_t* hr;
  const Tables& T;
  Sink& sink;
  Stream* rng;  // the current task's stream (external sampling swaps it per task)
  int up;
  float tw;  // strategy-sum weight (float)iteration, cfr_gpu.cu:828
  uint32_t path_hash = 2166136261u;
  int decisions = 0, updates = 0, terminal = 0, update_legal = 0, leaves = 0;
  std::vector<orc_decision>* log = nullptr;  // decision log (orc_walk_cards)

  void log_decision(const Game& g, int player, int a, uint32_t vm, uint32_t slot, uint32_t hist, int pot_before) {
    if (!log) return;
    orc_decision d{};
    d.player = (uint8_t)player;
    d.street = (uint8_t)g.street;
    d.action = (uint8_t)a;
    d.is_update = player == up;
    d.legal = vm;
    d.history = hist;
    d.slot = slot;
    d.chips = g.pot - pot_before;
    d.pot = g.pot;
    d.current_bet = g.current_bet;
    d.lfr = g.lfr;
    d.raises = g.raises;
    d.n_queue = g.n_to_act;
    for (int i = 0; i < g.n_to_act; ++i) d.queue |= (uint32_t)g.to_act[i] << (4 * i);
    log->push_back(d);
  }

  // Outcome sampling chains FNV-1a over the path's (slot, action) bytes in
  // walk order; external sampling adds each pair's FNV-1a (order free: the
  // device explores the tree's tasks in no fixed order).
  void hash_step(uint32_t slot, int a) {
    if (cfg.sampling == MCCFR_SAMPLE_EXTERNAL) {
      path_hash += fnv_pair(slot, a);
      return;
    }
    for (int i = 0; i < 4; ++i) path_hash = (path_hash ^ ((slot >> (8 * i)) & 0xFFu)) * 16777619u;
    path_hash = (path_hash ^ (uint32_t)(a & 0xFF)) * 16777619u;
  }

  // A leaf of kind 1 (fold), 2 (showdown) or 3 (depth guard): the trace keeps
  // the kind (outcome sampling: the path's one leaf) or the set of kinds
  // reached, bit k for kind k (external sampling).
  void leaf(int kind) {
    ++leaves;
    if (cfg.sampling == MCCFR_SAMPLE_EXTERNAL)
      terminal |= 1 << kind;
    else
      terminal = kind;
  }

  // cfr_gpu.cu:510-537
  int sample(uint32_t slot, uint32_t vm) {
    const int A = cfg.num_actions;
    float total = 0.0f;
    for (int a = 0; a < A; ++a)
      if (vm >> a & 1u) total += T.s(slot, a);
    if (total < 1e-7f) return passive_fallback(cfg, vm);
    const float r = u01(rng->next()) * total;
    float cum = 0.0f;
    int result = -1;
    for (int a = 0; a < A; ++a) {
      const bool valid = vm >> a & 1u;
      if (valid) cum += T.s(slot, a);
      if (valid && result < 0 && r <= cum) result = a;
    }
    if (result < 0) result = passive_fallback(cfg, vm);
    return result;
  }

  // cfr_gpu.cu:543-573 (+ raise counting for capped menus)
  bool apply(Game& g, int player, int a) {
    const int to_call = g.current_bet - g.bets[player];
    const int chips = chips_for(cfg, a, g.pot, g.stacks[player], to_call, g.current_bet, g.lfr);
    const uint32_t hmask = 0x3FFFFFFFu;
    g.ab = ((g.ab << cfg.history_bits) | (uint32_t)a) & hmask;
    const int kind = cfg.menu[a].kind;
    if (kind == MCCFR_ACT_FOLD) { g.folded |= 1u << player; return false; }
    if (kind == MCCFR_ACT_CHECK || (kind == MCCFR_ACT_CHECK_CALL && chips == 0)) return false;
    g.stacks[player] -= chips;
    g.bets[player] += chips;
    g.invested[player] += chips;
    g.pot += chips;
    if (g.stacks[player] == 0) g.all_in |= 1u << player;
    if (g.bets[player] > g.current_bet) {
      const int prev = g.current_bet;
      const int raise = g.bets[player] - prev;
      if (raise >= g.lfr || prev == 0) g.lfr = raise;
      g.current_bet = g.bets[player];
      ++g.raises;
      return true;
    }
    return false;
  }

  void rebuild_queue(Game& g, int raiser) {  // cfr_gpu.cu:575-585
    g.n_to_act = 0;
    for (int i = 1; i < g.N; ++i) {
      const int p = (raiser + i) % g.N;
      if (!(g.folded >> p & 1u) && !(g.all_in >> p & 1u)) g.to_act[g.n_to_act++] = p;
    }
  }

  float showdown(const Game& g) {  // cfr_gpu.cu:694-752
    uint16_t ranks[6] = {0, 0, 0, 0, 0, 0};
    for (int p = 0; p < g.N; ++p) {
      if (g.folded >> p & 1u) continue;
      const uint8_t c[7] = {g.hole[2 * p], g.hole[2 * p + 1], g.comm[0], g.comm[1],
                            g.comm[2], g.comm[3], g.comm[4]};
      ranks[p] = walk7(hr, c);
    }
    int levels[6], nL = 0;
    for (int p = 0; p < g.N; ++p) {
      const int inv = g.invested[p];
      if (inv == 0) continue;
      bool dup = false;
      for (int j = 0; j < nL; ++j) dup = dup || levels[j] == inv;
      if (!dup) levels[nL++] = inv;
    }
    std::sort(levels, levels + nL);
    float payoff = -(float)g.invested[up];
    int prev = 0;
    for (int li = 0; li < nL; ++li) {
      const int L = levels[li];
      int layer_pot = 0;
      for (int q = 0; q < g.N; ++q) layer_pot += std::max(0, std::min(g.invested[q], L) - prev);
      uint16_t best = 0;
      for (int q = 0; q < g.N; ++q)
        if (!(g.folded >> q & 1u) && g.invested[q] >= L) best = std::max(best, ranks[q]);
      if (best > 0) {
        int n_win = 0;
        for (int q = 0; q < g.N; ++q)
          if (!(g.folded >> q & 1u) && g.invested[q] >= L && ranks[q] == best) ++n_win;
        if (!(g.folded >> up & 1u) && g.invested[up] >= L && ranks[up] == best)
          payoff += (float)layer_pot / (float)n_win;
      }
      prev = L;
    }
    return payoff;
  }

  // cfr_gpu.cu:644-860.  Both branches follow exactly one child, so the
  // recursion is a single root-to-leaf path; the state is not restored after
  // the child returns because nothing reads it afterwards.
  float traverse(Game& g, int depth) {
    if (depth >= 20) { leaf(3); return 0.0f; }
    int active = 0, winner = -1;
    for (int i = 0; i < g.N; ++i)
      if (!(g.folded >> i & 1u)) { ++active; winner = i; }
    if (active == 1) {
      leaf(1);
      return winner == up ? (float)(g.pot - g.invested[up]) : (float)(-(int)g.invested[up]);
    }
    if (g.n_to_act == 0) {
      if (g.street < 3) {
        ++g.street;
        for (int i = 0; i < g.N; ++i) g.bets[i] = 0;
        g.current_bet = 0;
        g.lfr = g.bb_amt;
        g.raises = 0;
        g.n_to_act = 0;
        for (int i = 1; i <= g.N; ++i) {
          const int p = (g.dealer + i) % g.N;
          if (!(g.folded >> p & 1u) && !(g.all_in >> p & 1u)) g.to_act[g.n_to_act++] = p;
        }
        if (g.n_to_act > 0) return traverse(g, depth + 1);
      }
      leaf(2);
      return showdown(g);
    }
    const int player = g.to_act[0];
    const int to_call = g.current_bet - g.bets[player];
    const uint32_t vm = valid_mask(cfg, g.pot, g.stacks[player], to_call, g.current_bet, g.lfr, g.raises);
    const uint32_t slot = info_hash((uint8_t)player, g.hb[player], g.bbk[player][g.street],
                                    (uint8_t)g.street, g.ab, cfg.table_log2);
    const int A = cfg.num_actions;
    ++decisions;
    if (player == up) {
      ++updates;
      update_legal += __builtin_popcount(vm);
      float strat[MCCFR_MAX_ACTIONS];
      float sum_r = 0.0f;
      for (int a = 0; a < A; ++a) {
        const float r = (vm >> a & 1u) ? std::fmax(0.0f, T.r(slot, a)) : 0.0f;
        strat[a] = r;
        sum_r += r;
      }
      const int nvalid = __builtin_popcount(vm);
      for (int a = 0; a < A; ++a)
        strat[a] = (sum_r > 1e-7f) ? strat[a] / sum_r : ((vm >> a & 1u) ? 1.0f / (float)nvalid : 0.0f);
      if (cfg.sampling == MCCFR_SAMPLE_EXTERNAL) {
        // external sampling: every legal action's subtree, in action order,
        // from a copy of this node's state; no draw at this node.  Each
        // subtree is a task with its own stream: the traversal's counters
        // with the task's path code (es_child_path of this task's code and
        // the action), from draw 0.
        float v[MCCFR_MAX_ACTIONS] = {};
        Stream* const own = rng;
        for (int a = 0; a < A; ++a) {
          if (!(vm >> a & 1u)) continue;
          Game child = g;
          hash_step(slot, a);
          const uint32_t hist = child.ab;
          const int pot_before = child.pot;
          --child.n_to_act;
          for (int i = 0; i < child.n_to_act; ++i) child.to_act[i] = child.to_act[i + 1];
          if (apply(child, player, a)) rebuild_queue(child, player);
          log_decision(child, player, a, vm, slot, hist, pot_before);
          Stream task = *own;
          task.path = es_child_path(own->path, a);
          task.k = 0;
          task.blk_id = 0xFFFFFFFFu;
          rng = &task;
          v[a] = traverse(child, depth + 1);
          rng = own;
        }
        float vi = 0.0f;  // v(I) = sum_a sigma(a) v(a), in action order
        for (int a = 0; a < A; ++a)
          if (vm >> a & 1u) vi += strat[a] * v[a];
        for (int a = 0; a < A; ++a)
          if (vm >> a & 1u) sink.add(slot, a, v[a] - vi, tw * strat[a]);
        return vi;
      }
      const int sampled = sample(slot, vm);
      const float q = std::fmax(strat[sampled], 0.05f);
      hash_step(slot, sampled);
      const uint32_t hist = g.ab;
      const int pot_before = g.pot;
      --g.n_to_act;
      for (int i = 0; i < g.n_to_act; ++i) g.to_act[i] = g.to_act[i + 1];
      if (apply(g, player, sampled)) rebuild_queue(g, player);
      log_decision(g, player, sampled, vm, slot, hist, pot_before);
      const float child_ev = traverse(g, depth + 1);
      const float ev = child_ev;
      for (int a = 0; a < A; ++a) {
        if (!(vm >> a & 1u)) continue;
        const float cf_val = (a == sampled) ? (child_ev / q) : 0.0f;
        sink.add(slot, a, cf_val - ev, tw * strat[a]);
      }
      return ev;
    }
    const int chosen = sample(slot, vm);
    hash_step(slot, chosen);
    const uint32_t hist = g.ab;
    const int pot_before = g.pot;
    --g.n_to_act;
    for (int i = 0; i < g.n_to_act; ++i) g.to_act[i] = g.to_act[i + 1];
    if (apply(g, player, chosen)) rebuild_queue(g, player);
    log_decision(g, player, chosen, vm, slot, hist, pot_before);
    return traverse(g, depth + 1);
  }
};

// Deal + buckets of one traversal.  Preflop root (cfr_gpu.cu:921-937):
// 2N+5 draws over the full deck.  Postflop root (the project's re-solve root,
// mccfr.h mccfr_root_state): the fixed board cards leave the deck (the rest in
// ascending order), then 2N + 5 - fixed draws over the remaining cards.
void deal_cards(const mccfr_game_cfg& cfg, const mccfr_root_state* root, Stream& rng, Game& g) {
  const int N = cfg.num_players;
  const int nb = (root && root->street > 0) ? root->street + 2 : 0;
  uint8_t deck[52];
  int M = 0;
  for (int c = 0; c < 52; ++c) {
    bool fixed = false;
    for (int k = 0; k < nb; ++k) fixed = fixed || root->board[k] == c;
    if (!fixed) deck[M++] = (uint8_t)c;
  }
  const int n_deal = 2 * N + 5 - nb;
  for (int i = 0; i < n_deal; ++i) {
    const uint32_t r = rng.next();
    const int j = i + (int)(((uint64_t)r * (uint64_t)(M - i)) >> 32);
    std::swap(deck[i], deck[j]);
  }
  for (int i = 0; i < 2 * N; ++i) g.hole[i] = deck[i];
  for (int k = 0; k < 5; ++k) g.comm[k] = k < nb ? root->board[k] : deck[2 * N + k - nb];
}

// cfr_gpu.cu:921-981: deal, buckets, blinds and the preflop queue; or the
// postflop root: stacks and investments as given, seats outside active_mask
// folded, no bets, queue from the seat after the dealer (as at a street
// change, cfr_gpu.cu:676-686).
void setup_dealt(const mccfr_game_cfg& cfg, const mccfr_root_state* root, const int32_t* hr, const uint8_t* lut,
                 int preflop_dealer, Game& g);

void setup_game(const mccfr_game_cfg& cfg, const mccfr_root_state* root, const int32_t* hr, const uint8_t* lut,
                Stream& rng, Game& g) {
  deal_cards(cfg, root, rng, g);
  // draw 2N+5 of the preflop root (cfr_gpu.cu:963); a postflop root names its dealer
  const int dealer = (root && root->street > 0) ? root->dealer : (int)(rng.next() % (uint32_t)cfg.num_players);
  setup_dealt(cfg, root, hr, lut, dealer, g);
}

// Buckets, stacks, blinds and the first queue of a dealt hand (g.hole, g.comm set).
void setup_dealt(const mccfr_game_cfg& cfg, const mccfr_root_state* root, const int32_t* hr, const uint8_t* lut,
                 int preflop_dealer, Game& g) {
  const int N = cfg.num_players;
  g.N = N;
  for (int p = 0; p < N; ++p) g.hb[p] = lut[g.hole[2 * p] * 52 + g.hole[2 * p + 1]];
  for (int p = 0; p < N; ++p) {
    g.bbk[p][0] = 0;
    const uint8_t c[7] = {g.hole[2 * p], g.hole[2 * p + 1], g.comm[0], g.comm[1],
                          g.comm[2], g.comm[3], g.comm[4]};
    g.bbk[p][1] = bucket_of(walk5(hr, c));
    g.bbk[p][2] = bucket_of(walk6(hr, c));
    g.bbk[p][3] = bucket_of(walk7(hr, c));
  }
  g.current_bet = 0;
  g.lfr = cfg.big_blind;
  g.bb_amt = cfg.big_blind;
  g.raises = 0;
  g.n_to_act = 0;
  if (root && root->street > 0) {
    g.folded = 0;
    g.all_in = 0;
    g.pot = 0;
    for (int p = 0; p < N; ++p) {
      g.stacks[p] = root->stack[p];
      g.bets[p] = 0;
      g.invested[p] = root->invested[p];
      g.pot += root->invested[p];
      const bool act = root->active_mask >> p & 1u;
      if (!act) g.folded |= 1u << p;
      if (act && root->stack[p] == 0) g.all_in |= 1u << p;
    }
    g.ab = root->history & 0x3FFFFFFFu;
    g.street = root->street;
    g.dealer = root->dealer;
    for (int i = 1; i <= N; ++i) {
      const int p = (g.dealer + i) % N;
      if (!(g.folded >> p & 1u) && !(g.all_in >> p & 1u)) g.to_act[g.n_to_act++] = p;
    }
    return;
  }
  for (int i = 0; i < N; ++i) {
    g.stacks[i] = cfg.starting_stack;
    g.bets[i] = 0;
    g.invested[i] = 0;
  }
  g.folded = 0; g.all_in = 0; g.pot = 0;
  g.ab = 0; g.street = 0;
  g.dealer = preflop_dealer;
  const int sb_pos = (N == 2) ? g.dealer : (g.dealer + 1) % N;
  const int bb_pos = (N == 2) ? (g.dealer + 1) % N : (g.dealer + 2) % N;
  const int sb_amt = std::min(cfg.small_blind, g.stacks[sb_pos]);
  const int bb_amt = std::min(cfg.big_blind, g.stacks[bb_pos]);
  g.stacks[sb_pos] -= sb_amt; g.bets[sb_pos] = sb_amt; g.invested[sb_pos] = sb_amt;
  g.stacks[bb_pos] -= bb_amt; g.bets[bb_pos] = bb_amt; g.invested[bb_pos] = bb_amt;
  g.pot = sb_amt + bb_amt;
  g.current_bet = bb_amt;
  if (g.stacks[sb_pos] == 0) g.all_in |= 1u << sb_pos;
  if (g.stacks[bb_pos] == 0) g.all_in |= 1u << bb_pos;
  for (int i = 0; i < N; ++i) {
    const int p = (bb_pos + 1 + i) % N;
    if (!(g.all_in >> p & 1u)) g.to_act[g.n_to_act++] = p;
  }
}

template <class Tables, class Sink>
void run_traversal(const mccfr_game_cfg& cfg, const mccfr_root_state* root, const int32_t* hr, const uint8_t* lut,
                   const Tables& T, Sink& sink, int up, float tw, uint64_t tid, uint32_t tag, mccfr_trace* tr) {
  Stream rng(cfg.seed, tid, tag);
  Game g;
  setup_game(cfg, root, hr, lut, rng, g);
  Walker<Tables, Sink> w{cfg, hr, T, sink, &rng, up, tw};
  const float u = w.traverse(g, 0);
  if (tr) {
    tr->utility = u;
    tr->path_hash = w.path_hash;
    tr->decisions = (uint16_t)w.decisions;
    tr->updates = (uint8_t)w.updates;
    tr->terminal = (uint8_t)w.terminal;
    tr->update_legal = (uint16_t)w.update_legal;
    tr->reserved = (uint16_t)(cfg.sampling == MCCFR_SAMPLE_EXTERNAL ? w.leaves : 0);
  }
}

// ===========================================================================
// Strategy-quality harness (eval.cpp:45-627): seat 0 plays the average
// strategy, every other seat a scripted archetype, for n hands with the
// dealer button at hand % n.  Two random streams drive the same game:
//   EvalPhilox  the project's stream.  Hand h draws word k&3 of
//               Philox(counter = {k>>2, h_lo, h_hi, kEvalTag}, key = seed):
//               a fresh deck and 2n+5 partial Fisher-Yates draws (as the
//               traversal deal), uniforms (x>>8) * 2^-24 in [0, 1), and a
//               RANDOM opponent's pick x % count.  Hands are independent.
//   EvalMt      the reference's stream (eval.cpp:599-627, card.h:38-48): one
//               std::mt19937(seed) for the whole run, one Deck kept across
//               hands and reshuffled j = rng() % (i + 1) for i = 51..1,
//               std::uniform_real_distribution<float>(0, total).
// ===========================================================================
constexpr uint32_t kEvalTag = 0xE7A1u;
constexpr int kMaxEvalSeats = 9;

// abstraction.cpp:61-70: the CPU bucket clamps at 127 (the device training
// walk truncates to uint8 instead, cfr_gpu.cu:488; they differ only for rank 0).
inline int eval_bucket(uint16_t rank) {
  return std::min((int)((uint32_t)(rank - 1) * 128u / 7462u), 127);
}

struct EvalSpec {
  const mccfr_game_cfg* cfg;
  const int32_t* hr;
  const uint8_t* lut;
  const float* P;  // normalised strategy, slot-major rows of `row`
  int row;
  int n;
  int opp;
  int stack;
};

struct EvalHand {
  int n;
  uint8_t cards[2 * kMaxEvalSeats + 5];
  int stacks[kMaxEvalSeats], bets[kMaxEvalSeats], invested[kMaxEvalSeats];
  bool folded[kMaxEvalSeats], all_in[kMaxEvalSeats];
  int pot, current_bet, lfr, bb_amt, raises;
  uint32_t ab;
  int street;
  int hb[kMaxEvalSeats], bk[kMaxEvalSeats][4];
};

struct EvalPhilox {
  Stream s;
  EvalPhilox(uint64_t seed, uint64_t hand) : s(seed, hand, kEvalTag) {}
  void deal(int n, uint8_t* cards) {
    uint8_t deck[52];
    for (int i = 0; i < 52; ++i) deck[i] = (uint8_t)i;
    for (int i = 0; i < 2 * n + 5; ++i) {
      const uint32_t r = s.next();
      const int j = i + (int)(((uint64_t)r * (uint64_t)(52 - i)) >> 32);
      std::swap(deck[i], deck[j]);
    }
    std::memcpy(cards, deck, 2 * n + 5);
  }
  float uniform(float total) { return (float)(s.next() >> 8) * (1.0f / 16777216.0f) * total; }
  uint32_t pick(uint32_t count) { return s.next() % count; }
};

struct EvalMt {
  std::mt19937 rng;
  uint8_t deck[52];
  explicit EvalMt(unsigned seed) : rng(seed) {
    for (int i = 0; i < 52; ++i) deck[i] = (uint8_t)i;
  }
  void deal(int n, uint8_t* cards) {
    for (int i = 51; i > 0; --i) {
      const int j = (int)(rng() % (unsigned)(i + 1));
      std::swap(deck[i], deck[j]);
    }
    std::memcpy(cards, deck, 2 * n + 5);
  }
  float uniform(float total) { return std::uniform_real_distribution<float>(0.f, total)(rng); }
  uint32_t pick(uint32_t count) { return (uint32_t)(rng() % count); }
};

// eval.cpp:49-69
template <class Rng>
int eval_sample(const mccfr_game_cfg& g, const float* probs, uint32_t vm, Rng& rng) {
  const int A = g.num_actions;
  float total = 0.f;
  for (int a = 0; a < A; ++a)
    if (vm >> a & 1u) total += probs[a];
  if (total < 1e-7f) return passive_fallback(g, vm);
  const float r = rng.uniform(total);
  float cum = 0.f;
  int last = -1;
  for (int a = 0; a < A; ++a) {
    if (!(vm >> a & 1u)) continue;
    cum += probs[a];
    last = a;
    if (r <= cum) return a;
  }
  return last;
}

// eval.h:5-13 / eval.cpp:100-157: {fold, call, raise} per archetype; RANDOM = 4.
struct Archetype {
  float fold, call, raise;
};
constexpr Archetype kArchetypes[4] = {{0.08f, 0.72f, 0.20f}, {0.60f, 0.30f, 0.10f}, {0.10f, 0.20f, 0.70f},
                                      {0.28f, 0.44f, 0.28f}};

template <class Rng>
int eval_scripted(const mccfr_game_cfg& g, int opp, uint32_t vm, int to_call, Rng& rng) {
  const int A = g.num_actions;
  if (opp == 4) {
    const uint32_t k = rng.pick((uint32_t)__builtin_popcount(vm));
    uint32_t m = vm;
    for (uint32_t i = 0; i < k; ++i) m &= m - 1;
    return __builtin_ctz(m);
  }
  const Archetype& q = kArchetypes[opp];
  float probs[MCCFR_MAX_ACTIONS] = {};
  int check_a = -1, call_a = -1, fold_a = -1, n_raise = 0;
  for (int a = 0; a < A; ++a) {
    if (!(vm >> a & 1u)) continue;
    const int k = g.menu[a].kind;
    if (k == MCCFR_ACT_CHECK || (k == MCCFR_ACT_CHECK_CALL && to_call == 0)) {
      if (check_a < 0) check_a = a;
    } else if (k == MCCFR_ACT_CALL || k == MCCFR_ACT_CHECK_CALL) {
      if (call_a < 0) call_a = a;
    } else if (k == MCCFR_ACT_FOLD) {
      if (fold_a < 0) fold_a = a;
    } else {
      ++n_raise;
    }
  }
  if (check_a >= 0) {
    probs[check_a] = q.fold + q.call;
  } else {
    if (fold_a >= 0) probs[fold_a] = q.fold;
    if (call_a >= 0) probs[call_a] = q.call;
  }
  if (n_raise > 0) {
    const float per = q.raise / (float)n_raise;
    for (int a = 0; a < A; ++a) {
      const int k = g.menu[a].kind;
      if ((vm >> a & 1u) && (k == MCCFR_ACT_RAISE_MIN || k == MCCFR_ACT_POT_FRAC || k == MCCFR_ACT_ALL_IN))
        probs[a] = per;
    }
  } else if (check_a >= 0) {
    probs[check_a] += q.raise;
  } else if (call_a >= 0) {
    probs[call_a] += q.raise;
  } else if (fold_a >= 0) {
    probs[fold_a] += q.raise;
  }
  return eval_sample(g, probs, vm, rng);
}

// eval.cpp:85-97 with get_strategy's row selection (cfr_gpu.cu:1398-1401).
template <class Rng>
int eval_trained(const EvalSpec& E, const EvalHand& h, int p, uint32_t vm, Rng& rng) {
  const mccfr_game_cfg& g = *E.cfg;
  const uint32_t key = info_hash((uint8_t)p, (uint8_t)h.hb[p], (uint8_t)h.bk[p][h.street], (uint8_t)h.street, h.ab,
                                 g.table_log2);
  const float* row = E.P + (size_t)key * E.row;
  float t = 0.f;
  for (int a = 0; a < E.row; ++a) t += row[a];
  if (!(t >= 1e-6f)) return passive_fallback(g, vm);
  return eval_sample(g, row, vm, rng);
}

// Seat 0 plays the strategy, the others the archetype.
template <class Rng>
struct EvalPlayers {
  const EvalSpec& E;
  Rng& rng;
  bool bad = false;
  void deal(int n, uint8_t* cards) { rng.deal(n, cards); }
  int act(const EvalHand& h, int p, uint32_t vm, int to_call) {
    return p == 0 ? eval_trained(E, h, p, vm, rng) : eval_scripted(*E.cfg, E.opp, vm, to_call, rng);
  }
};

// Cards and every decision given (known-answer hands in the tests).
struct EvalScript {
  const uint8_t* cards;
  const int32_t* actions;
  int n_actions, used = 0;
  bool bad = false;
  void deal(int n, uint8_t* out) { std::memcpy(out, cards, 2 * n + 5); }
  int act(const EvalHand&, int, uint32_t vm, int) {
    if (used >= n_actions) { bad = true; return __builtin_ctz(vm); }
    const int a = actions[used++];
    if (a < 0 || a >= 32 || !(vm >> a & 1u)) { bad = true; return __builtin_ctz(vm); }
    return a;
  }
};

// eval.cpp:322-391
template <class Players>
bool eval_round(const EvalSpec& E, EvalHand& h, int first, Players& pl) {
  const mccfr_game_cfg& g = *E.cfg;
  const int n = h.n;
  bool need[kMaxEvalSeats] = {};
  for (int p = 0; p < n; ++p) need[p] = !h.folded[p] && !h.all_in[p];
  int p = first;
  const int safety = n * n * 4;
  for (int guard = 0; guard < safety; ++guard) {
    bool anyone = false;
    for (int i = 0; i < n; ++i) anyone = anyone || need[i];
    if (!anyone) break;
    if (!need[p]) { p = (p + 1) % n; continue; }
    const int to_call = h.current_bet - h.bets[p];
    const uint32_t vm = valid_mask(g, h.pot, h.stacks[p], to_call, h.current_bet, h.lfr, h.raises);
    const int chosen = pl.act(h, p, vm, to_call);
    h.ab = ((h.ab << g.history_bits) | (uint32_t)chosen) & 0x3FFFFFFFu;
    need[p] = false;
    const int kind = g.menu[chosen].kind;
    if (kind == MCCFR_ACT_FOLD) {
      h.folded[p] = true;
      int active = 0;
      for (int i = 0; i < n; ++i) active += !h.folded[i];
      if (active <= 1) return false;
    } else if (kind != MCCFR_ACT_CHECK) {
      const int chips = chips_for(g, chosen, h.pot, h.stacks[p], to_call, h.current_bet, h.lfr);
      const int prev = h.current_bet;
      h.stacks[p] -= chips;
      h.bets[p] += chips;
      h.invested[p] += chips;
      h.pot += chips;
      if (h.stacks[p] == 0) h.all_in[p] = true;
      if (h.bets[p] > h.current_bet) {
        const int raise = h.bets[p] - prev;
        if (raise >= h.lfr || prev == 0) h.lfr = raise;
        h.current_bet = h.bets[p];
        ++h.raises;
        for (int i = 0; i < n; ++i)
          if (i != p && !h.folded[i] && !h.all_in[i]) need[i] = true;
      }
    }
    p = (p + 1) % n;
  }
  return true;
}

// eval.cpp:479-572 plus the side-pot showdown of eval.cpp:410-466.  Writes
// every seat's net chips; a side pot no live seat contested (no eligible
// winner) is left unpaid.
template <class Players>
void eval_hand(const EvalSpec& E, int dealer, Players& pl, int32_t* net) {
  const mccfr_game_cfg& g = *E.cfg;
  EvalHand h;
  const int n = E.n;
  h.n = n;
  pl.deal(n, h.cards);
  for (int p = 0; p < n; ++p) {
    h.stacks[p] = E.stack;
    h.bets[p] = h.invested[p] = 0;
    h.folded[p] = h.all_in[p] = false;
  }
  h.pot = h.current_bet = 0;
  h.lfr = g.big_blind;
  h.bb_amt = g.big_blind;
  h.raises = 0;
  h.ab = 0;
  h.street = 0;
  const int sb_pos = n == 2 ? dealer : (dealer + 1) % n;
  const int bb_pos = n == 2 ? (dealer + 1) % n : (dealer + 2) % n;
  auto post = [&](int p, int amt) {
    const int chips = std::min(amt, h.stacks[p]);
    h.stacks[p] -= chips;
    h.bets[p] = chips;
    h.invested[p] = chips;
    h.pot += chips;
    if (!h.stacks[p]) h.all_in[p] = true;
  };
  post(sb_pos, g.small_blind);
  post(bb_pos, g.big_blind);
  h.current_bet = h.bets[bb_pos];
  const uint8_t* comm = h.cards + 2 * n;
  for (int p = 0; p < n; ++p) {
    const uint8_t c[7] = {h.cards[2 * p], h.cards[2 * p + 1], comm[0], comm[1], comm[2], comm[3], comm[4]};
    h.hb[p] = E.lut[c[0] * 52 + c[1]];
    h.bk[p][0] = 0;
    h.bk[p][1] = eval_bucket(walk5(E.hr, c));
    h.bk[p][2] = eval_bucket(walk6(E.hr, c));
    h.bk[p][3] = eval_bucket(walk7(E.hr, c));
  }
  for (int st = 0; st < 4; ++st) {
    h.street = st;
    if (st > 0) {
      for (int p = 0; p < n; ++p) h.bets[p] = 0;
      h.current_bet = 0;
      h.lfr = h.bb_amt;
      h.raises = 0;
    }
    int active = 0, can_act = 0;
    for (int p = 0; p < n; ++p) {
      active += !h.folded[p];
      can_act += !h.folded[p] && !h.all_in[p];
    }
    if (active <= 1) break;
    if (can_act <= 1) {
      if (st == 3) break;
      continue;
    }
    int first = st == 0 ? (bb_pos + 1) % n : (dealer + 1) % n;
    if (st > 0)
      for (int i = 0; i < n; ++i) {
        const int q = (dealer + 1 + i) % n;
        if (!h.folded[q] && !h.all_in[q]) { first = q; break; }
      }
    if (!eval_round(E, h, first, pl)) break;
    active = 0;
    for (int p = 0; p < n; ++p) active += !h.folded[p];
    if (active <= 1 || st == 3) break;
  }
  int last = -1, active = 0;
  for (int p = 0; p < n; ++p)
    if (!h.folded[p]) { last = p; ++active; }
  if (active == 1) {
    for (int p = 0; p < n; ++p) net[p] = (p == last ? h.pot : 0) - h.invested[p];
    return;
  }
  uint16_t ranks[kMaxEvalSeats];
  int win[kMaxEvalSeats], rem[kMaxEvalSeats];
  for (int p = 0; p < n; ++p) {
    const uint8_t c[7] = {h.cards[2 * p], h.cards[2 * p + 1], comm[0], comm[1], comm[2], comm[3], comm[4]};
    ranks[p] = h.folded[p] ? 0 : walk7(E.hr, c);
    win[p] = 0;
    rem[p] = h.invested[p];
  }
  while (true) {
    int lvl = INT32_MAX;
    for (int p = 0; p < n; ++p)
      if (rem[p] > 0 && rem[p] < lvl) lvl = rem[p];
    if (lvl == INT32_MAX) break;
    int side = 0;
    bool elig[kMaxEvalSeats] = {};
    for (int p = 0; p < n; ++p) {
      if (rem[p] <= 0) continue;
      const int c = std::min(rem[p], lvl);
      side += c;
      rem[p] -= c;
      if (!h.folded[p]) elig[p] = true;
    }
    uint16_t best = 0;
    for (int p = 0; p < n; ++p)
      if (elig[p] && ranks[p] > best) best = ranks[p];
    int nw = 0;
    for (int p = 0; p < n; ++p) nw += elig[p] && ranks[p] == best;
    if (nw == 0) continue;
    const int share = side / nw;
    for (int p = 0; p < n; ++p)
      if (elig[p] && ranks[p] == best) win[p] += share;
  }
  for (int p = 0; p < n; ++p) net[p] = win[p] - h.invested[p];
}

bool cfg_ok(const mccfr_game_cfg* c) {
  if (!c) return false;
  if (c->num_players < 2 || c->num_players > 6) return false;
  if (c->num_actions < 1 || c->num_actions > MCCFR_MAX_ACTIONS) return false;
  if (c->table_log2 < 4 || c->table_log2 > 32) return false;
  if (c->sampling != MCCFR_SAMPLE_OUTCOME && c->sampling != MCCFR_SAMPLE_EXTERNAL) return false;
  if (c->history_bits < 1 || c->history_bits > 8) return false;
  for (int a = 0; a < c->num_actions; ++a)
    if (c->menu[a].kind == MCCFR_ACT_POT_FRAC && c->menu[a].den == 0) return false;
  return true;
}

template <class F>
void parallel_for(int threads, int64_t n, F&& f) {
  if (threads <= 1 || n < 2) {
    f(0, (int64_t)0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    const int64_t b = n * t / threads, e = n * (t + 1) / threads;
    pool.emplace_back([&f, t, b, e]() { f(t, b, e); });
  }
  for (auto& th : pool) th.join();
}

}  // namespace

// ===========================================================================
// extern "C" surface
// ===========================================================================
extern "C" {

int orc_to_rm(int card) { return to_rm(card); }
uint16_t orc_decode_packed(uint16_t p) { return decode_packed(p); }
uint16_t orc_walk5(const int32_t* hr, const uint8_t* c) { return walk5(hr, c); }
uint16_t orc_walk6(const int32_t* hr, const uint8_t* c) { return walk6(hr, c); }
uint16_t orc_walk7(const int32_t* hr, const uint8_t* c) { return walk7(hr, c); }

void orc_walk7_batch(const int32_t* hr, const uint8_t* cards, size_t stride, uint16_t* out,
                     size_t n, int threads) {
  parallel_for(threads, (int64_t)n, [&](int, int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) out[i] = walk7(hr, cards + (size_t)i * stride);
  });
}

int orc_validate_table(const int32_t* hr, size_t words) {
  if (words != MCCFR_HAND_TABLE_WORDS) return 0;
  const uint8_t royal[7] = {51, 50, 49, 48, 47, 0, 14};  // As Ks Qs Js Ts 2c 3d
  return walk7(hr, royal) == 7462 ? 1 : 0;
}

int orc_bf_class_count(int per_category[9]) {
  const BFClasses& B = bf();
  for (int c = 0; c < 9; ++c) per_category[c] = B.per_cat[c];
  return B.count;
}

int orc_bf_class(const uint8_t* cards, int n) { return bf_class_n(cards, n); }

int orc_exhaustive5_bf(uint64_t cat_counts[10]) {
  for (int i = 0; i < 10; ++i) cat_counts[i] = 0;
  uint8_t c[5];
  for (c[0] = 0; c[0] < 52; ++c[0])
    for (c[1] = c[0] + 1; c[1] < 52; ++c[1])
      for (c[2] = c[1] + 1; c[2] < 52; ++c[2])
        for (c[3] = c[2] + 1; c[3] < 52; ++c[3])
          for (c[4] = c[3] + 1; c[4] < 52; ++c[4]) ++cat_counts[category_of_class(bf_class_n(c, 5))];
  return 0;
}

int orc_exhaustive7(const int32_t* hr, int threads, int check_bf, uint64_t cat_counts[10],
                    uint64_t* mismatches) {
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  if (check_bf) (void)bf();
  std::vector<std::array<uint64_t, 11>> local(threads);
  for (auto& l : local) l.fill(0);
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t]() {
      auto& acc = local[t];
      for (;;) {
        const int pair = next.fetch_add(1);
        if (pair >= 52 * 52) return;
        const int a = pair / 52, b = pair % 52;
        if (b <= a) continue;
        const int p2 = hr[hr[53 + to_rm(a)] + to_rm(b)];
        uint8_t c[7] = {(uint8_t)a, (uint8_t)b, 0, 0, 0, 0, 0};
        for (int x = b + 1; x < 52; ++x) {
          const int p3 = hr[p2 + to_rm(x)];
          c[2] = (uint8_t)x;
          for (int y = x + 1; y < 52; ++y) {
            const int p4 = hr[p3 + to_rm(y)];
            c[3] = (uint8_t)y;
            for (int z = y + 1; z < 52; ++z) {
              const int p5 = hr[p4 + to_rm(z)];
              c[4] = (uint8_t)z;
              for (int v = z + 1; v < 52; ++v) {
                const int p6 = hr[p5 + to_rm(v)];
                c[5] = (uint8_t)v;
                for (int w = v + 1; w < 52; ++w) {
                  const uint16_t r = decode_packed((uint16_t)hr[p6 + to_rm(w)]);
                  ++acc[category_of_class(r)];
                  if (check_bf) {
                    c[6] = (uint8_t)w;
                    if (bf_class_n(c, 7) != r) ++acc[10];
                  }
                }
              }
            }
          }
        }
      }
    });
  for (auto& th : pool) th.join();
  for (int i = 0; i < 10; ++i) cat_counts[i] = 0;
  uint64_t mm = 0;
  for (auto& l : local) {
    for (int i = 0; i < 10; ++i) cat_counts[i] += l[i];
    mm += l[10];
  }
  if (mismatches) *mismatches = mm;
  return 0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox(ctr, key, out);
}

uint32_t orc_draw_u32(uint64_t seed, uint64_t tid, uint32_t tag, uint32_t k) {
  const uint32_t ctr[4] = {k >> 2, (uint32_t)tid, (uint32_t)(tid >> 32), tag};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t out[4];
  philox(ctr, key, out);
  return out[k & 3];
}

float orc_u01(uint32_t x) { return u01(x); }
uint16_t orc_f32_to_bf16(float x) { return f2bf(x); }
float orc_bf16_to_f32(uint16_t h) { return bf2f(h); }

void orc_preflop_lut(uint8_t out[2704]) { preflop_lut(out); }

uint32_t orc_info_hash(uint8_t player, uint8_t hb, uint8_t bbk, uint8_t street, uint32_t ab,
                       int log2) {
  return info_hash(player, hb, bbk, street, ab, log2);
}

uint32_t orc_valid_mask(const mccfr_game_cfg* cfg, int pot, int stack, int to_call,
                        int current_bet, int lfr, int raises) {
  return valid_mask(*cfg, pot, stack, to_call, current_bet, lfr, raises);
}

int orc_chips(const mccfr_game_cfg* cfg, int a, int pot, int stack, int to_call, int current_bet,
              int lfr) {
  return chips_for(*cfg, a, pot, stack, to_call, current_bet, lfr);
}

int orc_passive_fallback(const mccfr_game_cfg* cfg, uint32_t mask) {
  return passive_fallback(*cfg, mask);
}

int orc_board_bucket(const int32_t* hr, const uint8_t* hole2, const uint8_t* comm5, int street) {
  const uint8_t c[7] = {hole2[0], hole2[1], comm5[0], comm5[1], comm5[2], comm5[3], comm5[4]};
  if (street == 0) return 0;
  if (street == 1) return bucket_of(walk5(hr, c));
  if (street == 2) return bucket_of(walk6(hr, c));
  return bucket_of(walk7(hr, c));
}

void orc_deal(uint64_t seed, uint64_t tid, uint32_t tag, int N, uint8_t* cards_out, int* dealer_out) {
  Stream rng(seed, tid, tag);
  uint8_t deck[52];
  for (int i = 0; i < 52; ++i) deck[i] = (uint8_t)i;
  for (int i = 0; i < 2 * N + 5; ++i) {
    const uint32_t r = rng.next();
    const int j = i + (int)(((uint64_t)r * (uint64_t)(52 - i)) >> 32);
    std::swap(deck[i], deck[j]);
  }
  for (int i = 0; i < 2 * N + 5; ++i) cards_out[i] = deck[i];
  *dealer_out = (int)(rng.next() % (uint32_t)N);
}

void orc_deal_root(const mccfr_game_cfg* cfg, const mccfr_root_state* root, uint64_t tid, uint32_t tag,
                   uint8_t* hole_out, uint8_t* board_out) {
  Stream rng(cfg->seed, tid, tag);
  Game g;
  g.N = cfg->num_players;
  deal_cards(*cfg, root, rng, g);
  for (int i = 0; i < 2 * g.N; ++i) hole_out[i] = g.hole[i];
  for (int k = 0; k < 5; ++k) board_out[k] = g.comm[k];
}

void orc_regret_matching(const float* R, uint16_t* sigma, uint64_t slots, int row) {
  for (uint64_t s = 0; s < slots; ++s) {
    const float* r = R + s * row;
    float sum = 0.0f;
    for (int a = 0; a < row; ++a) sum += std::fmax(0.0f, r[a]);
    for (int a = 0; a < row; ++a)
      sigma[s * row + a] = f2bf((sum > 1e-7f) ? std::fmax(0.0f, r[a]) / sum : 0.0f);
  }
}

int orc_replay_launch(const mccfr_game_cfg* cfg, const int32_t* hr, const uint8_t* lut,
                      const float* R, const uint16_t* sigma, int up, int64_t t_weight,
                      uint64_t first_tid, uint32_t n, uint32_t tag, double* dR, double* dS,
                      double* aR, double* aS, mccfr_trace* trace) {
  return orc_replay_launch_root(cfg, nullptr, hr, lut, R, sigma, up, t_weight, first_tid, n, tag, dR, dS, aR, aS,
                                trace);
}

int orc_replay_launch_root(const mccfr_game_cfg* cfg, const mccfr_root_state* root, const int32_t* hr,
                           const uint8_t* lut, const float* R, const uint16_t* sigma, int up, int64_t t_weight,
                           uint64_t first_tid, uint32_t n, uint32_t tag, double* dR, double* dS, double* aR,
                           double* aS, mccfr_trace* trace) {
  if (!cfg_ok(cfg) || !hr || !lut || !R) return -1;
  const bool accumulate = dR && dS && aR && aS;
  if (!accumulate && !trace) return -1;
  if (up < 0 || up >= cfg->num_players) return -1;
  const int row = row_width(*cfg);
  const uint64_t slots = 1ull << cfg->table_log2;
  std::vector<uint16_t> own;
  if (!sigma) {
    own.resize(slots * row);
    orc_regret_matching(R, own.data(), slots, row);
    sigma = own.data();
  }
  TablesAoS T{R, sigma, row};
  const float tw = (float)t_weight;
  if (accumulate) {
    SinkF64 sink{dR, dS, aR, aS, row};
    for (uint32_t i = 0; i < n; ++i)
      run_traversal(*cfg, root, hr, lut, T, sink, up, tw, first_tid + i, tag, trace ? trace + i : nullptr);
  } else {
    SinkNull sink;
    for (uint32_t i = 0; i < n; ++i)
      run_traversal(*cfg, root, hr, lut, T, sink, up, tw, first_tid + i, tag, trace + i);
  }
  return 0;
}

int64_t orc_replay_sparse(const mccfr_game_cfg* cfg, const mccfr_root_state* root, const int32_t* hr,
                          const uint8_t* lut, const float* R, const uint16_t* sigma, int up, int64_t t_weight,
                          uint64_t first_tid, uint32_t n, uint32_t tag, int threads, mccfr_trace* trace,
                          orc_sparse* out) {
  if (!cfg_ok(cfg) || !hr || !lut || !R || !sigma || !out) return -1;
  if (up < 0 || up >= cfg->num_players) return -1;
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  const int row = row_width(*cfg);
  TablesAoS T{R, sigma, row};
  const float tw = (float)t_weight;
  std::vector<std::vector<Contribution>> lists(threads);
  parallel_for(threads, (int64_t)n, [&](int th, int64_t b, int64_t e) {
    SinkList sink{&lists[th]};
    for (int64_t i = b; i < e; ++i)
      run_traversal(*cfg, root, hr, lut, T, sink, up, tw, first_tid + (uint64_t)i, tag, trace ? trace + i : nullptr);
  });
  // concatenate in traversal order (thread t holds an ascending range), then a
  // stable sort by table index keeps each entry's contributions in that order
  std::vector<Contribution> all;
  size_t total = 0;
  for (auto& l : lists) total += l.size();
  all.reserve(total);
  for (auto& l : lists) {
    all.insert(all.end(), l.begin(), l.end());
    std::vector<Contribution>().swap(l);
  }
  auto idx_of = [row](const Contribution& c) { return (uint64_t)c.slot * (uint64_t)row + c.a; };
  std::stable_sort(all.begin(), all.end(),
                   [&](const Contribution& x, const Contribution& y) { return idx_of(x) < idx_of(y); });
  size_t uniq = 0;
  for (size_t i = 0; i < all.size(); ++i) uniq += (i == 0 || idx_of(all[i]) != idx_of(all[i - 1]));
  out->n = (int64_t)uniq;
  out->idx = static_cast<uint64_t*>(std::malloc(std::max<size_t>(1, uniq) * sizeof(uint64_t)));
  double** cols[4] = {&out->dR, &out->dS, &out->aR, &out->aS};
  for (double** c : cols) *c = static_cast<double*>(std::malloc(std::max<size_t>(1, uniq) * sizeof(double)));
  if (!out->idx || !out->dR || !out->dS || !out->aR || !out->aS) {
    orc_sparse_free(out);
    return -1;
  }
  int64_t k = -1;
  for (size_t i = 0; i < all.size(); ++i) {
    const uint64_t x = idx_of(all[i]);
    if (i == 0 || x != idx_of(all[i - 1])) {
      ++k;
      out->idx[k] = x;
      out->dR[k] = out->dS[k] = out->aR[k] = out->aS[k] = 0.0;
    }
    out->dR[k] += (double)all[i].rd;
    out->dS[k] += (double)all[i].sd;
    out->aR[k] += std::fabs((double)all[i].rd);
    out->aS[k] += std::fabs((double)all[i].sd);
  }
  return out->n;
}

void orc_sparse_free(orc_sparse* s) {
  if (!s) return;
  std::free(s->idx);
  std::free(s->dR);
  std::free(s->dS);
  std::free(s->aR);
  std::free(s->aS);
  s->idx = nullptr;
  s->dR = s->dS = s->aR = s->aS = nullptr;
  s->n = 0;
}

int orc_walk_cards(const mccfr_game_cfg* cfg, const mccfr_root_state* root, const int32_t* hr, const uint8_t* lut,
                   const uint8_t* hole, const uint8_t* board, int dealer, const float* R, const uint16_t* sigma,
                   int up, int64_t t_weight, uint64_t tid, uint32_t tag, mccfr_trace* trace, orc_decision* log,
                   int log_cap, int* n_log, uint32_t* c_slot, uint32_t* c_action, float* c_rd, float* c_sd,
                   int c_cap, int* n_contrib) {
  if (!cfg_ok(cfg) || !hr || !lut || !R || !hole || !board) return -1;
  const int N = cfg->num_players;
  if (up < 0 || up >= N || dealer < 0 || dealer >= N) return -1;
  const int row = row_width(*cfg);
  const uint64_t slots = 1ull << cfg->table_log2;
  std::vector<uint16_t> own;
  if (!sigma) {
    own.resize(slots * row);
    orc_regret_matching(R, own.data(), slots, row);
    sigma = own.data();
  }
  TablesAoS T{R, sigma, row};
  Game g;
  for (int i = 0; i < 2 * N; ++i) g.hole[i] = hole[i];
  for (int k = 0; k < 5; ++k) g.comm[k] = board[k];
  const bool postflop = root && root->street > 0;
  setup_dealt(*cfg, root, hr, lut, postflop ? root->dealer : dealer, g);
  // the walk's draws follow the deal's (2N + 5 - fixed cards, then the
  // preflop dealer draw), as in a replayed launch
  Stream rng(cfg->seed, tid, tag);
  const int n_draws = 2 * N + 5 - (postflop ? root->street + 2 : 0) + (postflop ? 0 : 1);
  for (int i = 0; i < n_draws; ++i) (void)rng.next();
  std::vector<Contribution> contrib;
  SinkList sink{&contrib};
  std::vector<orc_decision> dec;
  Walker<TablesAoS, SinkList> w{*cfg, hr, T, sink, &rng, up, (float)t_weight};
  w.log = &dec;
  const float u = w.traverse(g, 0);
  if (trace) {
    trace->utility = u;
    trace->path_hash = w.path_hash;
    trace->decisions = (uint16_t)w.decisions;
    trace->updates = (uint8_t)w.updates;
    trace->terminal = (uint8_t)w.terminal;
    trace->update_legal = (uint16_t)w.update_legal;
    trace->reserved = (uint16_t)(cfg->sampling == MCCFR_SAMPLE_EXTERNAL ? w.leaves : 0);
  }
  if (n_log) *n_log = (int)dec.size();
  if (n_contrib) *n_contrib = (int)contrib.size();
  if (log) {
    if ((int)dec.size() > log_cap) return -1;
    std::copy(dec.begin(), dec.end(), log);
  }
  if (c_slot || c_action || c_rd || c_sd) {
    if ((int)contrib.size() > c_cap) return -1;
    for (size_t i = 0; i < contrib.size(); ++i) {
      if (c_slot) c_slot[i] = contrib[i].slot;
      if (c_action) c_action[i] = contrib[i].a;
      if (c_rd) c_rd[i] = contrib[i].rd;
      if (c_sd) c_sd[i] = contrib[i].sd;
    }
  }
  return 0;
}

double orc_cpu_baseline(const mccfr_game_cfg* cfg, const int32_t* hr, const uint8_t* lut,
                        float* R, float* S, uint16_t* sig, int launches, uint32_t batch,
                        int first_player, int64_t first_iter, uint64_t first_tid, int threads) {
  if (!cfg_ok(cfg) || cfg->num_actions > 8) return -1.0;
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  const uint64_t slots = 1ull << cfg->table_log2;
  const int A = 8;
  std::vector<std::vector<Contribution>> lists(threads);
  for (auto& l : lists) l.reserve((size_t)batch / threads * 16 + 1024);
  const auto t0 = std::chrono::steady_clock::now();
  int player = first_player;
  int64_t iter = first_iter;
  uint64_t tid = first_tid;
  for (int L = 0; L < launches; ++L) {
    // kernel_regret_matching (cfr_gpu.cu:318-351) over the whole SoA table
    parallel_for(threads, (int64_t)slots, [&](int, int64_t b, int64_t e) {
      for (int64_t s = b; s < e; ++s) {
        float sum = 0.0f;
        for (int a = 0; a < A; ++a) sum += std::fmax(0.0f, R[(size_t)a * slots + s]);
        for (int a = 0; a < A; ++a)
          sig[(size_t)a * slots + s] =
              f2bf((sum > 1e-7f) ? std::fmax(0.0f, R[(size_t)a * slots + s]) / sum : 0.0f);
      }
    });
    // kernel_simulate_batch (cfr_gpu.cu:878-1011), traversals split over threads
    TablesSoA T{R, sig, slots};
    const float tw = cfg->use_linear_cfr ? (float)iter : 1.0f;
    parallel_for(threads, (int64_t)batch, [&](int t, int64_t b, int64_t e) {
      lists[t].clear();
      SinkList sink{&lists[t]};
      for (int64_t i = b; i < e; ++i)
        run_traversal(*cfg, nullptr, hr, lut, T, sink, player, tw, tid + (uint64_t)i, 0u, nullptr);
    });
    for (auto& l : lists)
      for (const Contribution& c : l) {
        R[(size_t)c.a * slots + c.slot] += c.rd;
        S[(size_t)c.a * slots + c.slot] += c.sd;
      }
    // kernel_clamp_regrets (cfr_gpu.cu:1018-1022)
    if (cfg->use_cfr_plus)
      parallel_for(threads, (int64_t)(slots * A), [&](int, int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) R[i] = std::fmax(0.0f, R[i]);
      });
    tid += batch;
    if (++player == cfg->num_players) { player = 0; ++iter; }
  }
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// ---- strategy-quality harness (eval.cpp:581-627) ----
int orc_eval_hands(const mccfr_game_cfg* cfg, const int32_t* hr, const uint8_t* lut, const float* P, int row,
                   int n_players, int opponent, int stack, uint64_t seed, uint64_t first_hand, int64_t n_hands,
                   int stream, int threads, int32_t* bot_net, int32_t* seat_net) {
  if (!cfg || !hr || !lut || !P || n_players < 2 || n_players > kMaxEvalSeats || opponent < 0 || opponent > 4 ||
      n_hands < 0 || (stream == 1 && first_hand != 0) || stream < 0 || stream > 1)
    return -1;
  const EvalSpec E{cfg, hr, lut, P, row, n_players, opponent, stack};
  const int n = n_players;
  auto store = [&](int64_t i, const int32_t* net) {
    if (bot_net) bot_net[i] = net[0];
    if (seat_net)
      for (int p = 0; p < n; ++p) seat_net[i * n + p] = net[p];
  };
  if (stream == 1) {
    EvalMt rng((unsigned)seed);
    EvalPlayers<EvalMt> pl{E, rng};
    for (int64_t h = 0; h < n_hands; ++h) {
      int32_t net[kMaxEvalSeats];
      eval_hand(E, (int)(h % n), pl, net);
      store(h, net);
    }
    return 0;
  }
  parallel_for(threads, n_hands, [&](int, int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const uint64_t h = first_hand + (uint64_t)i;
      EvalPhilox rng(seed, h);
      EvalPlayers<EvalPhilox> pl{E, rng};
      int32_t net[kMaxEvalSeats];
      eval_hand(E, (int)(h % (uint64_t)n), pl, net);
      store(i, net);
    }
  });
  return 0;
}

int orc_eval_scripted(const mccfr_game_cfg* cfg, const int32_t* hr, const uint8_t* lut, int n_players, int stack,
                      int dealer, const uint8_t* cards, const int32_t* actions, int n_actions, int32_t* seat_net) {
  if (!cfg || n_players < 2 || n_players > kMaxEvalSeats) return -1;
  const EvalSpec E{cfg, hr, lut, nullptr, 0, n_players, 0, stack};
  EvalScript pl{cards, actions, n_actions};
  eval_hand(E, dealer, pl, seat_net);
  if (pl.bad) return -2;
  return pl.used;
}

int orc_eval_scripted_action(const mccfr_game_cfg* cfg, int opponent, uint32_t vm, int to_call, uint64_t seed,
                             int64_t draws, int64_t* counts) {
  if (!cfg || opponent < 0 || opponent > 4 || !vm) return -1;
  for (int a = 0; a < MCCFR_MAX_ACTIONS; ++a) counts[a] = 0;
  for (int64_t i = 0; i < draws; ++i) {
    EvalPhilox rng(seed, (uint64_t)i);
    ++counts[eval_scripted(*cfg, opponent, vm, to_call, rng)];
  }
  return 0;
}

int orc_eval_buckets(const int32_t* hr, const uint8_t* lut, int n_players, const uint8_t* cards, int32_t* hb,
                     int32_t* bk) {
  const uint8_t* comm = cards + 2 * n_players;
  for (int p = 0; p < n_players; ++p) {
    const uint8_t c[7] = {cards[2 * p], cards[2 * p + 1], comm[0], comm[1], comm[2], comm[3], comm[4]};
    hb[p] = lut[c[0] * 52 + c[1]];
    bk[4 * p + 0] = 0;
    bk[4 * p + 1] = eval_bucket(walk5(hr, c));
    bk[4 * p + 2] = eval_bucket(walk6(hr, c));
    bk[4 * p + 3] = eval_bucket(walk7(hr, c));
  }
  return 0;
}

}  // extern "C"
