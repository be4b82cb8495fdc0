// Launch wrappers of the gfx950 kernels (defined in kernels.hip and traverse.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace mccfr {

// Update-player nodes a lane keeps in LDS during its walk; deeper ones are
// written to the lane's node list in HBM when they are reached.
constexpr int kRegPending = 4;

// Radix sort of the launch's records: ceil(sort_bits / 10) passes with the
// sorted bits spread evenly over them (26 bits: three 9-bit digits), tiles of
// kSortWG threads x kSortItems keys.
constexpr int kSortWG = 256;
constexpr int kSortItems = 8;
constexpr uint32_t kSortTile = kSortWG * kSortItems;
constexpr uint32_t kDigitBits = 10;     // widest radix digit (a launch spreads its key bits evenly over passes)
constexpr uint32_t kMinDigitBits = 8;   // narrowest (a scatter thread owns whole digits)
constexpr uint32_t kSortBins = 1u << kDigitBits;

// External sampling by generations (one-segment launches): generation 0's
// tasks are the traversals' roots; a task walks (sampling the other seats'
// actions) to a leaf or to an update-player node, and such a node makes one
// task of the next generation per legal action.  Two update-player decisions
// on a path are at least two levels apart (another seat acts between them,
// or a street change takes a level), so under the depth guard (decisions at
// levels 0..19) a path holds at most 10 update nodes: tasks of generations
// 0..10.  The last generation flags a node it would make.
constexpr uint32_t kEsGens = 11;
// Nodes and tasks are allocated in kEsRegions regions, each with its own
// counters on a line of its own (a wave uses region wave % kEsRegions), since
// atomics on one address serialize across the device.
constexpr uint32_t kEsRegions = 64;
constexpr uint32_t kEsCtrStride = 32;  // dwords between two counters (128 bytes)
// Buffers for a scratch of cap_lanes lanes: a node per record of the sort
// input (cap_node = 20 per lane), and as many tasks per generation.
struct EsGenBufs {
  uint32_t* task[2];   // [kEsRegions][cap_task_r] tasks of generation >= 1: node << 4 | action, by parity
  uint32_t* task_cnt;  // [kEsGens + 1][kEsRegions][kEsCtrStride] tasks of each generation and region
  uint32_t* node_cnt;  // [kEsGens + 1][kEsRegions][kEsCtrStride] nodes each generation made in each region
  uint4* node;         // [kEsRegions][cap_node_r][4] a node's state before its action, its slot and legal mask
  float* root_v;       // [cap_lanes] root value of each traversal (trace launches)
  uint32_t* tr_acc;    // [cap_lanes][8] trace sums per traversal (trace launches)
  uint32_t cap_task_r, cap_node_r;  // per region
};
constexpr uint32_t kEsNoParent = 0xFFFFFFFFu;

// Scratch of one launch, sized for `cap_lanes` lanes.  Node lists are
// lane-strided: entry k of lane l at [k * cap_lanes + l].
struct DevScratch {
  uint32_t* node_slot;   // [kMaxDepth][lanes] info-set slot of each update-player node
  uint32_t* node_av;     // [kMaxDepth][lanes] sampled action | legal mask << 8
  uint8_t* node_count;   // [lanes] update-player nodes of the traversal (0 on padding lanes)
  float* node_u;         // [lanes] terminal utility of the update player
  uint32_t* root_bkt;    // [kMaxPlayers][lanes] hole | flop << 8 | turn << 16 | river << 24 buckets
                         // (turn = 0xFF until the walk fetches it; real buckets are <= 198)
  uint16_t* root_rank;   // [kMaxPlayers][lanes] 7-card rank (showdown)
  uint16_t* sd_inv;      // [kMaxPlayers][lanes] chips each seat put in, for traversals ending in a showdown
  uint8_t* sd_state;     // [lanes] 0x80 | folded seats after a showdown, else 0
  uint32_t* root_tidx;   // [kMaxPlayers][lanes] colex rank of hole + first four board cards:
                         // the turn bucket's index, read only when the seat acts on the turn
  const uint8_t* turn6;  // the turn bucket table (set for each walk launch)
  uint8_t* root_dealer;  // [lanes]
  uint32_t* blk_cnt;     // [lanes / kEmitLanes] records per emission block, then their offsets
  uint32_t* key[2];      // [kMaxDepth * lanes] record keys (segment << slot_bits | slot), ping-pong
  uint2* val[2];         // [kMaxDepth * lanes] record payloads: (action | legal mask << 8, utility bits)
  uint32_t* hist;        // [kSortBins][max_tiles] tile digit counts, then per-digit prefixes over tiles
  uint32_t* bin_total;   // [kSortBins] keys per digit in the current pass
  // Per-wave regions of the sorted records (64 records each): region w lists
  // its short runs at runs[64 w ..] and the runs it defers at cand[4 w ..].
  uint4* runs;           // [max_waves][64] short runs: first record, records, key
  uint32_t* n_runs;      // [max_waves] short runs of each region
  uint2* cand;           // [max_waves][4] deferred runs: first record, end (0: past the region)
  uint32_t* n_cand;      // [max_waves] deferred runs of each region
  uint8_t* head0;        // [max_waves + kHeadPad] lane of each region's first run head (64: none)
  uint8_t* head1;        // [head1_bytes] per 64 regions: 0 if one of them holds a run head, else 64
  SegCtx* seg_ctx;       // [kMaxSegs] table pointers and weights of the launch's segments
  uint4* long_runs;      // [max_long] long runs: first record, records, first piece, pieces
  uint4* piece;          // [max_pieces] pieces of long runs: first record, end, key, long run
  double* partial;       // [max_pieces][2 * 16] FP64 sums of each piece
  LaunchCounters* cnt;
  uint32_t* es_stack;    // external-sampling frames [kEsFrames][words][launch lanes] (not in the arena)
  float* es_rd;          // external sampling: [sort input][row] regret deltas of each node record (not in the arena);
                         // the generation walk first collects each node's child values there
  EsGenBufs gen;         // external sampling by generations (not in the arena)
  uint32_t cap_lanes;
  uint32_t max_tiles;
};
constexpr uint32_t kEmitLanes = 256;   // lanes per emission block (one per thread)
constexpr uint32_t kShortRun = 16;     // longer runs are split into pieces reduced by whole workgroups
constexpr uint32_t kPiece = 256;       // records per piece of a long run (one wave each)
constexpr uint32_t kHeadPad = 128;     // head0 / head1 bytes past the last entry (whole 64-byte reads)
// Size of head1 for `cap_lanes` lanes, in whole dwords.
__host__ __device__ constexpr uint32_t head1_bytes(uint32_t cap_lanes) {
  const uint64_t regions = (uint64_t)cap_lanes * kMaxDepth / 64 + 1;
  return (uint32_t)((regions / 64 + 1 + kHeadPad + 3) & ~3ull);
}

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

// Shared, read-only inputs of the traversal.
struct DealInputs {
  BucketTables bt;
  ArithTables at;
  const uint8_t* lut;  // [52][52] preflop buckets
};

// Shape of the walk launch (autotuned): workgroup size and traversals per lane.
struct WalkShape {
  uint32_t wg = 256;
  uint32_t tpl = 1;
};

void launch_build_buckets(hipStream_t s, const int32_t* hr, BucketTables b, bool clamp = false);
void launch_set_tenant(hipStream_t s, DevGame* dst, const DevGame& g);

// One launch (traverse.hip): deal and walk of every segment's traversals
// (launch_walk), then records -> sort -> owner-computes table update
// (launch_reduce, whose stable radix passes order the records by the low
// `sort_bits` bits of their keys).  `trace` (device, indexed by lane) may be null.
// after_deal / after_walk (may be null): recorded after k_deal and after k_walk
// (before the emission), for per-kernel timing.
void launch_walk(hipStream_t s, int row, const LaunchCfg& lc, const DevSeg* segs, const DevGame* tenants,
                 const DealInputs& in, const DevScratch& sc, WalkShape shape, mccfr_trace* trace,
                 hipEvent_t after_deal = nullptr, hipEvent_t after_walk = nullptr);
// advance (may be null): the launch's segment table, moved on to the next
// launch's segments by run discovery (k_runs) once the emission has read it
// (a captured training launch: one kernel fewer per replay).
void launch_reduce(hipStream_t s, int row, const LaunchCfg& lc, const DevSeg* segs, const DevGame* tenants,
                   const DevScratch& sc, uint32_t sort_bits, DevSeg* advance = nullptr);

// Bytes of external-sampling walk frames for a launch of `lanes` lanes.
size_t es_stack_bytes(int row, uint32_t lanes);
// Bytes of the generation walk's buffers for a scratch of `cap_lanes` lanes,
// and their carving.
size_t es_gen_bytes(uint32_t cap_lanes);
void carve_es_gen(EsGenBufs& b, char* base, uint32_t cap_lanes);

void launch_rm_sweep(hipStream_t s, int row, const float* R, uint16_t* sig, uint64_t slots);
void launch_normalize(hipStream_t s, int row, const float* S, float* out, uint64_t slots);
void launch_zero(hipStream_t s, void* p, size_t bytes);
void launch_clamp(hipStream_t s, int row, float* R, uint64_t slots);
// plane: a tenant's regret or strategy-sum pointer (rows at stride 2 * row)
void launch_plane(hipStream_t s, int row, float* plane, float* dense, uint64_t slot0, uint64_t n, bool to_dense);
void launch_eval7_walk(hipStream_t s, const int32_t* hr, const uint8_t* cards, size_t stride,
                       uint16_t* out, size_t n);
void launch_eval7_arith(hipStream_t s, ArithTables t, const uint8_t* cards, size_t stride,
                        uint16_t* out, size_t n);
void launch_random_hands(hipStream_t s, uint64_t seed, uint32_t tag, uint8_t* out, size_t n);
void launch_randomize(hipStream_t s, int row, int num_actions, float* R, float* S, uint64_t slots, uint64_t seed);
void launch_copy(hipStream_t s, const uint4* src, uint4* dst, size_t n16);
void launch_gather(hipStream_t s, const uint4* table, uint64_t rows, uint32_t* out, size_t n);

}  // namespace mccfr
