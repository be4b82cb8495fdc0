// Launch wrappers of the gfx950 kernels (defined in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace mccfr {

// Update-player nodes a lane keeps in registers; deeper ones go to the
// lane's overflow entries in HBM.
constexpr int kRegPending = 4;

// Device tables of one game (slot-major rows of `row` entries) and the
// per-launch node lists (lane-strided: entry k of lane l at [k * lanes + l]).
struct DevTables {
  float* R;             // regrets
  float* S;             // strategy sums
  uint16_t* sig;        // BF16 regret-matched strategy snapshot
  double* D;            // per-launch FP64 deltas: [slot][0..row) regret, [row..2row) strategy sum
  uint32_t* stamp;      // last launch epoch whose update pass handled the slot
  uint32_t* node_slot;  // [kMaxDepth][lanes] info-set slot of each update-player node
  uint32_t* node_av;    // [kMaxDepth][lanes] sampled action | legal mask << 8
  uint8_t* node_count;  // [lanes] update-player nodes of the traversal
  float* node_u;        // [lanes] terminal utility of the update player
  uint32_t* root_bkt;   // [kMaxPlayers][lanes] hole | flop << 8 | turn << 16 | river << 24 buckets
  uint16_t* root_rank;  // [kMaxPlayers][lanes] 7-card rank (showdown)
  uint8_t* root_dealer; // [lanes]
  uint32_t* upd_list;   // [kMaxDepth * lanes] slots to update (one entry per slot per scatter tile)
  uint32_t lanes;       // lane capacity of the lists (max batch)
};

// Tables of the arithmetic 7-card evaluator (built on the host).
constexpr int kNonFlushClasses = 49205;  // rank multisets of 7 cards, <= 4 per rank
struct ArithTables {
  uint16_t* nonflush;  // [49205]
  uint16_t* flush;     // [8192]
  uint16_t* offsets;   // [13][8][5]
};

// Flop and turn buckets of every 5- and 6-card set, indexed by colex rank.
constexpr uint32_t kFlopSets = 2598960u;   // C(52, 5)
constexpr uint32_t kTurnSets = 20358520u;  // C(52, 6)
struct BucketTables {
  uint8_t* flop5;  // [kFlopSets]
  uint8_t* turn6;  // [kTurnSets]
};

void launch_build_buckets(hipStream_t s, const int32_t* hr, BucketTables b);
void launch_set_state(hipStream_t s, LaunchState* st, int update_player, int64_t iteration,
                      int64_t t_weight, uint64_t first_tid, uint32_t n);
void launch_advance_state(hipStream_t s, LaunchState* st, int num_players, int use_linear_cfr,
                          uint32_t batch);
// Deal + walk + scatter of one launch (traverse.hip).  walk_blocks = number of
// persistent 256-lane walk workgroups (0: one lane per traversal); scatter_wg =
// scatter workgroup size (256, 512 or 1024 threads, 4 traversals per thread).
void launch_traverse(hipStream_t s, int row, const DevGame* g, LaunchState* st, BucketTables bt,
                     ArithTables at, const uint8_t* lut, DevTables t, uint32_t n_lanes, uint32_t walk_blocks,
                     uint32_t scatter_wg, mccfr_trace* trace);
void launch_update(hipStream_t s, int row, const DevGame* g, LaunchState* st, DevTables t, uint32_t n_lanes);
void launch_rm_sweep(hipStream_t s, int row, DevTables t, uint64_t slots);
void launch_normalize(hipStream_t s, int row, const float* S, float* out, uint64_t slots);
void launch_eval7_walk(hipStream_t s, const int32_t* hr, const uint8_t* cards, size_t stride,
                       uint16_t* out, size_t n);
void launch_eval7_arith(hipStream_t s, ArithTables t, const uint8_t* cards, size_t stride,
                        uint16_t* out, size_t n);
void launch_random_hands(hipStream_t s, uint64_t seed, uint32_t tag, uint8_t* out, size_t n);
void launch_randomize(hipStream_t s, int row, int num_actions, DevTables t, uint64_t slots, uint64_t seed);
void launch_copy(hipStream_t s, const uint4* src, uint4* dst, size_t n16);

}  // namespace mccfr
