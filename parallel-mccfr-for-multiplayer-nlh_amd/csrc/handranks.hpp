// Rank-walk hand table ("HandRanks.dat" layout) and the 7462-class hand ranking.
//
// The table is the classic Two Plus Two state machine that the reference
// loads from disk (poker_cuda/src/hand_eval.cpp:12-14, cfr_gpu.cu:1086-1104):
// 32,487,834 int32 words = 53 words for an unused header block plus one
// 53-word block per canonical card-set state.  Block i of state i starts at
// 53*i + 53; word (base + c) for a rank-major card c in 1..52 holds either the
// base of the successor state (fewer than 7 cards) or the packed rank of the
// completed 7-card hand; word (base + 0) of a 5- or 6-card state holds that
// state's own packed rank.  Packed ranks are (category << 12) | ordinal with
// category 1 (high card) .. 9 (straight flush), which the reference decodes
// with fixed class offsets (hand_eval.cpp:29-45).
#pragma once
#include <cstddef>
#include <cstdint>

namespace mccfr {

constexpr size_t kHandTableWords = 32487834u;
constexpr int kHandStates = 612977;  // canonical states incl. the empty root
constexpr int kNumHandClasses = 7462;

// Fills `hr` (kHandTableWords words).  Uses `threads` host threads (0 = all).
// Returns the number of canonical states generated (kHandStates when correct).
int generate_hand_table(int32_t* hr, int threads);

// Hand strength of the best 5-card hand among n (5..7) cards given as
// suit-major codes 0..51 (card = suit*13 + rank).  1 = worst high card,
// 7462 = royal flush.
int hand_class(const uint8_t* cards, int n);

// Packs a 1..7462 class the way the table stores it: (category << 12) | ordinal.
uint32_t pack_hand_class(int cls);

}  // namespace mccfr
