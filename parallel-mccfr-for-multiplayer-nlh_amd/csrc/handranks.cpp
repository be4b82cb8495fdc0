// Generator for the rank-walk hand table and the 7462-class hand ranking.
//
// The reference ships only a stub generator (poker_cuda/tools/gen_lookup_table.cpp:63-96)
// and expects the classic Two Plus Two "HandRanks.dat" file (hand_eval.cpp:12-14).
// The trainer's flop and turn buckets read the *state pointer* left by a 5- or
// 6-card walk (cfr_gpu.cu:114-122, 140-150), so the state numbering must be
// exactly the classic one.  That numbering is fixed by the published
// state-machine algorithm: a state is the set of cards seen so far, written as
// one byte per card ((rank+1) << 4 | suit nibble), sorted in descending byte
// order and packed little-endian into a 64-bit id; a suit nibble is cleared
// (0) once that suit can no longer reach five cards; states are numbered by
// ascending id (the empty set is state 0).  This file restates that algorithm
// from its description; no code is shared with any other implementation.
#include "handranks.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

namespace mccfr {
namespace {

// ---------------------------------------------------------------------------
// 7462-class ranking.  A hand's strength key orders hands exactly like poker:
// key = category << 20 | five 4-bit tie-break ranks (rank 0 = deuce .. 12 = ace).
// Categories: 0 high card, 1 pair, 2 two pair, 3 trips, 4 straight, 5 flush,
// 6 full house, 7 quads, 8 straight flush.
// ---------------------------------------------------------------------------
inline uint32_t make_key(int cat, int k1 = 0, int k2 = 0, int k3 = 0, int k4 = 0, int k5 = 0) {
  return (uint32_t)cat << 20 | (uint32_t)k1 << 16 | (uint32_t)k2 << 12 | (uint32_t)k3 << 8 |
         (uint32_t)k4 << 4 | (uint32_t)k5;
}

// Highest straight in a 13-bit rank mask: top rank, 3 for the wheel, -1 if none.
inline int straight_top(uint32_t m) {
  for (int top = 12; top >= 4; --top)
    if (((m >> (top - 4)) & 0x1Fu) == 0x1Fu) return top;
  if ((m & 0x100Fu) == 0x100Fu) return 3;
  return -1;
}

// The n highest set ranks of m, skipping ranks in `skip`.
inline void top_ranks(uint32_t m, uint32_t skip, int n, int* out) {
  int k = 0;
  for (int r = 12; r >= 0 && k < n; --r)
    if ((m >> r & 1u) && !(skip >> r & 1u)) out[k++] = r;
  while (k < n) out[k++] = 0;
}

uint32_t key_flush(uint32_t m) {
  const int sf = straight_top(m);
  if (sf >= 0) return make_key(8, sf);
  int k[5];
  top_ranks(m, 0, 5, k);
  return make_key(5, k[0], k[1], k[2], k[3], k[4]);
}

uint32_t key_ranks(const int cnt[13]) {
  uint32_t present = 0;
  int quad = -1, trips[3], npairs = 0, ntrips = 0, pairs[4];
  for (int r = 12; r >= 0; --r) {
    if (cnt[r] >= 1) present |= 1u << r;
    if (cnt[r] == 4 && quad < 0) quad = r;
    else if (cnt[r] == 3) trips[ntrips++] = r;
    else if (cnt[r] == 2) pairs[npairs++] = r;
  }
  int k[5];
  if (quad >= 0) {
    top_ranks(present, 1u << quad, 1, k);
    return make_key(7, quad, k[0]);
  }
  if (ntrips >= 1 && (ntrips >= 2 || npairs >= 1)) {
    int p = -1;
    if (ntrips >= 2) p = trips[1];
    if (npairs >= 1 && pairs[0] > p) p = pairs[0];
    return make_key(6, trips[0], p);
  }
  const int st = straight_top(present);
  if (st >= 0) return make_key(4, st);
  if (ntrips == 1) {
    top_ranks(present, 1u << trips[0], 2, k);
    return make_key(3, trips[0], k[0], k[1]);
  }
  if (npairs >= 2) {
    top_ranks(present, (1u << pairs[0]) | (1u << pairs[1]), 1, k);
    return make_key(2, pairs[0], pairs[1], k[0]);
  }
  if (npairs == 1) {
    top_ranks(present, 1u << pairs[0], 3, k);
    return make_key(1, pairs[0], k[0], k[1], k[2]);
  }
  top_ranks(present, 0, 5, k);
  return make_key(0, k[0], k[1], k[2], k[3], k[4]);
}

// key -> class 1..7462, built once by enumerating every 5-card rank pattern.
struct ClassMap {
  std::vector<uint16_t> cls;  // indexed by key (< 9 << 20)
  int distinct = 0;
  ClassMap() : cls(9u << 20, 0) {
    std::vector<uint32_t> keys;
    keys.reserve(8000);
    int cnt[13];
    // every multiset of 5 ranks with at most 4 copies of a rank
    for (int a = 0; a < 13; ++a)
      for (int b = a; b < 13; ++b)
        for (int c = b; c < 13; ++c)
          for (int d = c; d < 13; ++d)
            for (int e = d; e < 13; ++e) {
              std::memset(cnt, 0, sizeof cnt);
              ++cnt[a]; ++cnt[b]; ++cnt[c]; ++cnt[d]; ++cnt[e];
              bool ok = true;
              for (int r = 0; r < 13; ++r) ok = ok && cnt[r] <= 4;
              if (!ok) continue;
              keys.push_back(key_ranks(cnt));
              if (a != b && b != c && c != d && d != e)
                keys.push_back(key_flush((1u << a) | (1u << b) | (1u << c) | (1u << d) | (1u << e)));
            }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    distinct = (int)keys.size();
    for (size_t i = 0; i < keys.size(); ++i) cls[keys[i]] = (uint16_t)(i + 1);
  }
};

const ClassMap& class_map() {
  static const ClassMap m;
  return m;
}

// Class of a card set given as (rank, suit) with suit -1 meaning "this card's
// suit cannot contribute to a flush".
int class_of(const int* rank, const int* suit, int n) {
  int cnt[13] = {0};
  uint32_t smask[4] = {0, 0, 0, 0};
  int scnt[4] = {0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    ++cnt[rank[i]];
    if (suit[i] >= 0) {
      smask[suit[i]] |= 1u << rank[i];
      ++scnt[suit[i]];
    }
  }
  uint32_t key = key_ranks(cnt);
  for (int s = 0; s < 4; ++s)
    if (scnt[s] >= 5) key = std::max(key, key_flush(smask[s]));
  return class_map().cls[key];
}

// ---------------------------------------------------------------------------
// Canonical card-set states.
// ---------------------------------------------------------------------------
struct Next {
  uint64_t id;  // 0 = rejected (repeated card or five of a rank)
  int ncards;   // cards in the set including the new one (also when rejected)
};

// Adds rank-major card c (1..52: c-1 = rank*4 + suit) to state `id`.
inline Next add_card(uint64_t id, int c) {
  int b[7];
  const int z = c - 1;
  b[0] = (((z >> 2) + 1) << 4) | ((z & 3) + 1);
  int n = 1;
  for (int i = 0; i < 6; ++i) {
    const int x = (int)((id >> (8 * i)) & 0xFFu);
    if (!x) break;
    b[n++] = x;
  }
  int scount[5] = {0, 0, 0, 0, 0};
  int rcount[14] = {0};
  bool repeat = false;
  for (int i = 0; i < n; ++i) {
    ++scount[b[i] & 0xF];
    ++rcount[(b[i] >> 4) & 0xF];
    if (i && b[i] == b[0]) repeat = true;
  }
  if (repeat) return {0, n};
  if (n > 4)
    for (int r = 1; r <= 13; ++r)
      if (rcount[r] > 4) return {0, n};
  const int need = n - 2;  // cards of one suit still needed for a flush to be possible
  if (need > 1)
    for (int i = 0; i < n; ++i)
      if (scount[b[i] & 0xF] < need) b[i] &= 0xF0;
  std::sort(b, b + n, [](int x, int y) { return x > y; });
  uint64_t out = 0;
  for (int i = 0; i < n; ++i) out |= (uint64_t)b[i] << (8 * i);
  return {out, n};
}

inline int state_cards(uint64_t id) {
  int n = 0;
  while (n < 8 && ((id >> (8 * n)) & 0xFFu)) ++n;
  return n;
}

uint32_t packed_class_of_state(uint64_t id) {
  int rank[7], suit[7];
  const int n = state_cards(id);
  for (int i = 0; i < n; ++i) {
    const int b = (int)((id >> (8 * i)) & 0xFFu);
    rank[i] = (b >> 4) - 1;
    suit[i] = (b & 0xF) ? (b & 0xF) - 1 : -1;
  }
  return pack_hand_class(class_of(rank, suit, n));
}

}  // namespace

uint32_t pack_hand_class(int cls) {
  // class boundaries of the 9 categories (counts 1277/2860/858/858/10/1277/156/156/10)
  static const int lo[10] = {0, 1277, 4137, 4995, 5853, 5863, 7140, 7296, 7452, 7462};
  for (int c = 0; c < 9; ++c)
    if (cls <= lo[c + 1]) return (uint32_t)((c + 1) << 12) | (uint32_t)(cls - lo[c]);
  return 0;
}

int hand_class(const uint8_t* cards, int n) {
  int rank[7], suit[7];
  for (int i = 0; i < n; ++i) {
    rank[i] = cards[i] % 13;
    suit[i] = cards[i] / 13;
  }
  return class_of(rank, suit, n);
}

int generate_hand_table(int32_t* hr, int threads) {
  if (class_map().distinct != kNumHandClasses) return -1;
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());

  // 1. All reachable states, level by level (a state with k cards has a larger
  //    id than any state with fewer cards, so concatenated levels are sorted).
  std::vector<uint64_t> states{0};
  std::vector<uint64_t> level{0};
  for (int k = 0; k < 6; ++k) {
    std::vector<uint64_t> next;
    next.reserve(level.size() * 52);
    for (uint64_t s : level)
      for (int c = 1; c <= 52; ++c) {
        const Next a = add_card(s, c);
        if (a.id && a.ncards < 7) next.push_back(a.id);
      }
    std::sort(next.begin(), next.end());
    next.erase(std::unique(next.begin(), next.end()), next.end());
    states.insert(states.end(), next.begin(), next.end());
    level.swap(next);
  }
  const int nstates = (int)states.size();
  if ((size_t)53 * ((size_t)nstates + 1) != kHandTableWords) return nstates;

  auto index_of = [&](uint64_t id) -> int {
    if (!id) return 0;
    return (int)(std::lower_bound(states.begin(), states.end(), id) - states.begin());
  };

  // 2. Transition words, threaded over states.
  std::memset(hr, 0, kHandTableWords * sizeof(int32_t));
  std::atomic<int> next_chunk{0};
  const int chunk = 4096;
  auto worker = [&]() {
    for (;;) {
      const int begin = next_chunk.fetch_add(chunk);
      if (begin >= nstates) return;
      const int end = std::min(nstates, begin + chunk);
      for (int i = begin; i < end; ++i) {
        const uint64_t s = states[i];
        const size_t base = (size_t)53 * i + 53;
        const int have = state_cards(s);
        for (int c = 1; c <= 52; ++c) {
          const Next a = add_card(s, c);
          if (a.ncards < 7)
            hr[base + c] = 53 * index_of(a.id) + 53;
          else
            hr[base + c] = a.id ? (int32_t)packed_class_of_state(a.id) : 0;
        }
        if (have == 5 || have == 6) hr[base] = (int32_t)packed_class_of_state(s);
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  return nstates;
}

}  // namespace mccfr
