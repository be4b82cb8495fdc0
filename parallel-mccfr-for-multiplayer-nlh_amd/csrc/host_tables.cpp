// Host-built lookup tables uploaded to the device: the preflop bucket table
// and the arithmetic 7-card evaluator's tables.
#include <algorithm>
#include <cstring>
#include <vector>

#include "handranks.hpp"
#include "host_tables.hpp"

namespace mccfr {
namespace {

// Chen-formula variant of abstraction.cpp:10-37 (rank 0 = deuce .. 12 = ace;
// base = rank/2 below the jack, straight bonus only below the ace).
float chen(int c0, int c1) {
  int hi = c0 % 13, lo = c1 % 13;
  if (hi < lo) std::swap(hi, lo);
  const bool suited = c0 / 13 == c1 / 13;
  float base = hi == 12 ? 10.f : hi == 11 ? 8.f : hi == 10 ? 7.f : hi == 9 ? 6.f : std::max(0.5f, hi / 2.f);
  if (hi == lo) return std::max(5.f, base * 2.f);
  float s = base + (suited ? 2.f : 0.f);
  const int gap = hi - lo - 1;
  static const float gap_pen[5] = {0.f, 1.f, 2.f, 4.f, 5.f};
  s -= gap_pen[gap < 4 ? gap : 4];
  if (gap <= 1 && hi < 12) s += 1.f;
  return s;
}

// ways[i][k]: count vectors over ranks i..12 summing to k with each count <= 4.
struct Ways {
  uint32_t w[14][8];
  Ways() {
    std::memset(w, 0, sizeof w);
    w[13][0] = 1;
    for (int i = 12; i >= 0; --i)
      for (int k = 0; k < 8; ++k)
        for (int c = 0; c <= 4 && c <= k; ++c) w[i][k] += w[i + 1][k - c];
  }
};

}  // namespace

// abstraction.cpp:42-59 (percentile of the score among all 1,326 hole pairs,
// 64 buckets) and cfr_gpu.cu:1164-1173 (table layout, zero diagonal).
void build_preflop_lut(uint8_t* lut) {
  std::vector<float> scores;
  scores.reserve(1326);
  for (int i = 0; i < 52; ++i)
    for (int j = i + 1; j < 52; ++j) scores.push_back(chen(i, j));
  std::sort(scores.begin(), scores.end());
  for (int a = 0; a < 52; ++a)
    for (int b = 0; b < 52; ++b) {
      uint8_t v = 0;
      if (a != b) {
        const float s = chen(a, b);
        const size_t below = (size_t)(std::lower_bound(scores.begin(), scores.end(), s) - scores.begin());
        const float pct = (float)below / (float)scores.size();
        v = (uint8_t)std::min((int)(pct * 64), 63);
      }
      lut[a * 52 + b] = v;
    }
}

void build_arith_tables(std::vector<uint16_t>& nonflush, std::vector<uint16_t>& flush,
                        std::vector<uint16_t>& offsets) {
  static const Ways W;
  // offsets[(i*8 + k)*5 + c] = sum_{c' < c} ways[i+1][k - c']
  offsets.assign(13 * 8 * 5, 0);
  for (int i = 0; i < 13; ++i)
    for (int k = 0; k < 8; ++k)
      for (int c = 0; c <= 4; ++c) {
        uint32_t acc = 0;
        for (int cc = 0; cc < c && cc <= k; ++cc) acc += W.w[i + 1][k - cc];
        offsets[(i * 8 + k) * 5 + c] = (uint16_t)acc;
      }
  // non-flush classes of every 7-card rank multiset
  nonflush.assign(W.w[0][7] + 1, 0);
  int cnt[13];
  uint8_t cards[7];
  auto visit = [&](auto&& self, int r, int left) -> void {
    if (r == 13) {
      if (left) return;
      uint32_t idx = 0;
      int k = 7, t = 0;
      for (int i = 0; i < 13; ++i) {
        idx += offsets[(i * 8 + k) * 5 + cnt[i]];
        k -= cnt[i];
        for (int c = 0; c < cnt[i]; ++c, ++t) cards[t] = (uint8_t)((t % 4) * 13 + i);  // suits cycle: no flush
      }
      nonflush[idx] = (uint16_t)hand_class(cards, 7);
      return;
    }
    for (int c = 0; c <= 4 && c <= left; ++c) {
      cnt[r] = c;
      self(self, r + 1, left - c);
    }
  };
  visit(visit, 0, 7);
  // flush / straight-flush class of a suit holding the ranks of `m` (5..7 bits)
  flush.assign(8192, 0);
  for (uint32_t m = 0; m < 8192; ++m) {
    const int bits = __builtin_popcount(m);
    if (bits < 5 || bits > 7) continue;
    int t = 0;
    for (int r = 0; r < 13; ++r)
      if (m >> r & 1u) cards[t++] = (uint8_t)r;  // all clubs
    flush[m] = (uint16_t)hand_class(cards, t);
  }
}

}  // namespace mccfr
