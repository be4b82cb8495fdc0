/*
 * mccfr.h — C ABI of the MI355X MCCFR trainer for multiplayer no-limit hold'em.
 *
 * Every entry point is plain C: integers, pointers and sizes; no C++ or HIP
 * types cross the boundary.  Each call returns 0 (MCCFR_OK) on success or a
 * negative MCCFR_E* code; mccfr_last_error() gives the text.  The library
 * never exits the process.  A context owns one HIP stream and all device
 * buffers it allocates; a context is not thread-safe (one host thread at a time).
 *
 * The drop-in target is the reference trainer class GPUCFRTrainer
 * (poker_cuda/src/cfr_gpu.cuh:48-132).  include/mccfr/gpu_cfr_trainer.hpp maps
 * that class onto these calls with the reference's signatures and defaults.
 */
#ifndef MCCFR_MI355X_H
#define MCCFR_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCCFR_OK 0
#define MCCFR_E_ARG (-1)      /* invalid argument or configuration */
#define MCCFR_E_HIP (-2)      /* HIP runtime error (message has the HIP text) */
#define MCCFR_E_TABLE (-3)    /* hand table missing, wrong size or failed the royal-flush check */
#define MCCFR_E_IO (-4)       /* file could not be opened, read or written */
#define MCCFR_E_FORMAT (-5)   /* file has a wrong magic, table size or action count */
#define MCCFR_E_NOMEM (-6)    /* device allocation failed */
#define MCCFR_E_STATE (-7)    /* call not valid in the context's current state */

#define MCCFR_MAX_PLAYERS 6
#define MCCFR_MAX_ACTIONS 16
#define MCCFR_HAND_TABLE_WORDS 32487834u /* poker_cuda/src/hand_eval.cpp:14 */

/* Action kinds of a betting menu.  The reference's fixed 8-action menu
 * (poker_cuda/src/abstraction.h:42-53) is
 *   {FOLD, CHECK, CALL, RAISE_MIN, POT_FRAC 1/2, POT_FRAC 2/3, POT_FRAC 1/1, ALL_IN}. */
enum mccfr_action_kind {
  MCCFR_ACT_FOLD = 0,
  MCCFR_ACT_CHECK = 1,      /* legal iff to_call == 0 */
  MCCFR_ACT_CALL = 2,       /* legal iff 0 < to_call < stack */
  MCCFR_ACT_CHECK_CALL = 3, /* check when to_call == 0, call when 0 < to_call < stack */
  MCCFR_ACT_RAISE_MIN = 4,  /* raise to the minimum legal total */
  MCCFR_ACT_POT_FRAC = 5,   /* raise to current_bet + max(1, num*pot/den), floored at min raise */
  MCCFR_ACT_ALL_IN = 6
};

typedef struct mccfr_action {
  uint8_t kind; /* enum mccfr_action_kind */
  uint8_t num;  /* POT_FRAC numerator */
  uint8_t den;  /* POT_FRAC denominator (> 0) */
  uint8_t reserved;
} mccfr_action;

/* Game and abstraction: the constructor arguments of GPUCFRTrainer
 * (cfr_gpu.cuh:50-55) plus the knobs the reference hard-codes. */
typedef struct mccfr_game_cfg {
  int32_t num_players;      /* 2..6 */
  int32_t starting_stack;   /* chips per seat (reference default 1000), at most 32767 */
  int32_t small_blind;      /* reference default 10 */
  int32_t big_blind;        /* reference default 20 */
  int32_t use_cfr_plus;     /* clamp regrets at 0 after every launch */
  int32_t use_linear_cfr;   /* strategy-sum weight t = iteration (else 1) */
  int32_t num_actions;      /* menu length, 1..16 */
  int32_t max_raises;       /* aggressive actions allowed per street; 0 = unlimited (reference) */
  int32_t history_bits;     /* bits appended to the action history per action (reference: 3) */
  int32_t table_log2;       /* info-set slots = 2^table_log2 (reference: 26) */
  uint64_t seed;            /* Philox key of the traversal streams */
  mccfr_action menu[MCCFR_MAX_ACTIONS];
  int32_t sampling;         /* MCCFR_SAMPLE_OUTCOME or MCCFR_SAMPLE_EXTERNAL (below) */
  int32_t reserved;
} mccfr_game_cfg;

/* Sampling at the update player's nodes.
 * OUTCOME (0, the reference's, cfr_gpu.cu:768-838): one action is sampled
 *   from the strategy snapshot and weighted by 1 / max(sigma, 0.05); a
 *   traversal is one root-to-leaf path.
 * EXTERNAL (1, textbook external sampling; no counterpart in the reference,
 *   which labels itself external sampling at cfr_gpu.cu:2 and :627-630 but
 *   samples): every legal action is explored, in action order; other seats
 *   sample one action from the snapshot as in OUTCOME.  A node's value is
 *   v(I) = sum_a sigma(a) v(a) with sigma from the FP32 regrets (as the
 *   reference's update node forms it), its regret contributions are
 *   v(a) - v(I) for every legal a, its strategy-sum contributions t * sigma(a)
 *   (the reference's rule, cfr_gpu.cu:828-836).  The tree is split into
 *   tasks: the root's, and one per legal action of every update node (the
 *   subtree below that action).  A task draws from its own stream: word k of
 *   Philox(counter = {k >> 2 | path << 6, traversal id, tag}, key = seed),
 *   path 0 for the root task (whose draws follow the deal's), and for the
 *   task under action a of a node in a task with code p,
 *   path = ((p ^ (a + 1)) * 16777619) mod 2^26, or 1 where that is 0.
 *   An update node draws nothing.  A traversal's utility is its root value. */
#define MCCFR_SAMPLE_OUTCOME 0
#define MCCFR_SAMPLE_EXTERNAL 1

/* Fills `cfg` with the reference game: its 8-action menu, no raise cap,
 * 3-bit history and 2^26 slots (cfr_gpu.cuh:17-25, abstraction.h:42-53). */
void mccfr_game_cfg_reference(mccfr_game_cfg* cfg, int num_players, int starting_stack,
                              int small_blind, int big_blind, int use_cfr_plus,
                              int use_linear_cfr);

/* Per-traversal record written by mccfr_run_launch when a trace buffer is
 * given (test and debugging aid; costs one 16-byte store per traversal). */
typedef struct mccfr_trace {
  float utility;       /* terminal payoff of the update player (EXTERNAL: the root value) */
  uint32_t path_hash;  /* FNV-1a over (slot, action) of every decision on the path
                          (EXTERNAL: 2166136261 plus, mod 2^32, the FNV-1a of each
                          (node, action) entered, in any order) */
  uint16_t decisions;  /* decision nodes on the path (EXTERNAL: in the explored tree) */
  uint8_t updates;     /* update-player decision nodes on the path (EXTERNAL: in the tree) */
  uint8_t terminal;    /* 1 fold, 2 showdown, 3 depth limit (EXTERNAL: bit k set when a
                          leaf of kind k was reached) */
  uint16_t update_legal; /* legal actions summed over the update-player nodes */
  uint16_t reserved;     /* EXTERNAL: leaves of the explored tree */
} mccfr_trace;         /* EXTERNAL: counts are taken modulo their field widths */

typedef struct mccfr_ctx mccfr_ctx;

const char* mccfr_last_error(const mccfr_ctx* ctx); /* ctx may be NULL (global errors) */

/* Creates a context on the current HIP device.  Allocates the FP32 regret and
 * strategy-sum tables, the BF16 strategy snapshot and the launch scratch for
 * `max_batch` traversals.  Replaces GPUCFRTrainer::GPUCFRTrainer +
 * alloc_device_buffers (cfr_gpu.cu:1064-1136). */
int mccfr_create(const mccfr_game_cfg* cfg, uint32_t max_batch, mccfr_ctx** out);
void mccfr_destroy(mccfr_ctx* ctx);

/* Hand table.  mccfr_generate_hand_table writes the 32,487,834-word rank-walk
 * table into a caller buffer (host code, no device needed).  mccfr_set_hand_table
 * validates a host table (size + royal-flush check, hand_eval.cpp:59-86) and
 * copies it to HBM; mccfr_load_hand_table does the same from a file
 * (GPUCFRTrainer::load_hand_table, cfr_gpu.cu:1086-1104);
 * mccfr_use_generated_hand_table generates it in host memory and uploads it. */
int mccfr_generate_hand_table(int32_t* hr, size_t words, int threads);
int mccfr_set_hand_table(mccfr_ctx* ctx, const int32_t* hr, size_t words);
int mccfr_load_hand_table(mccfr_ctx* ctx, const char* path);
int mccfr_use_generated_hand_table(mccfr_ctx* ctx);

/* Host-only hand-table checks (no context, no device), the ones
 * mccfr_set_hand_table and mccfr_load_hand_table apply before any device work.
 * The reference's loader checks the size and the royal flush
 * (hand_eval.cpp:59-86, cfr_gpu.cu:270-299); these also reject a table with
 * any word outside [0, words - 53], so that no walk over an accepted table
 * can read outside it.  The file form checks the file size before it
 * allocates.  MCCFR_OK, MCCFR_E_TABLE (wrong size, word out of range, royal
 * flush not 7462), MCCFR_E_IO (cannot open or read), MCCFR_E_ARG (null). */
int mccfr_validate_hand_table(const int32_t* hr, size_t words);
int mccfr_validate_hand_table_file(const char* path);

/* Preflop bucket table (52*52 bytes, [h0*52 + h1]) the trainer uses
 * (cfr_gpu.cu:1164-1173 over abstraction.cpp:10-59). */
int mccfr_preflop_lut(uint8_t* out_2704);

/* Seven-card ranks 1..7462 of n hands (7 suit-major card codes each, rows of
 * `stride` bytes, stride >= 7) already in device memory; ranks written to
 * device memory.  evaluator 0 = rank-walk table in HBM (eval7_gpu,
 * cfr_gpu.cu:126-137), 1 = arithmetic perfect-hash evaluator (no table).
 * A row with a code above 51 or a repeated card gets rank 0. */
int mccfr_eval7_device(mccfr_ctx* ctx, const uint8_t* d_cards, size_t stride, uint16_t* d_ranks,
                       size_t n, int evaluator);
/* Same over host buffers (copies in and out; for tests). */
int mccfr_eval7_host(mccfr_ctx* ctx, const uint8_t* cards, size_t stride, uint16_t* ranks,
                     size_t n, int evaluator);

/* One launch of the hot path: traversals first_tid .. first_tid+n-1 for
 * `update_player` with strategy-sum weight `t_weight`, then the touched-slot
 * regret update (CFR+ clamp) and strategy refresh.  This is one replay of the
 * reference graph [regret_matching -> simulate_batch -> clamp]
 * (cfr_gpu.cu:1286-1322).  `trace` is a HOST array of n records or NULL. */
int mccfr_run_launch(mccfr_ctx* ctx, int update_player, int64_t t_weight, uint64_t first_tid,
                     uint32_t n, mccfr_trace* trace);

/* mccfr_run_launch without the trace, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The launch runs on a
 * static device arena of the library (about 24 MB, held by each device the
 * library runs on until the process ends) instead of the context's launch
 * scratch, and calls no hipMalloc, hipFree or other allocating routine: n is
 * 1..16384 (MCCFR_E_ARG beyond).  Same kernels, same results as
 * mccfr_run_launch.  Every call on one device shares that device's arena, so
 * calls that overlap on one device (contexts on different streams) race:
 * serialize them.  The kernels initialise every arena word they read, so the
 * previous call's contents need no clearing. */
int mccfr_run_launch_noalloc(mccfr_ctx* ctx, int update_player, int64_t t_weight, uint64_t first_tid,
                             uint32_t n);

/* The reference training loop (GPUCFRTrainer::train, cfr_gpu.cu:1244-1378):
 * for iter = 1..iters, for p = 0..N-1, one launch of `batch` traversals,
 * replayed from a captured hipGraph.  Traversal ids continue from the
 * context's running counter. */
int mccfr_train(mccfr_ctx* ctx, int64_t iters, uint32_t batch);

/* Grows (or shrinks) the per-launch scratch to `max_batch` traversals; the
 * tables are kept.  Needed when a later train() uses a larger batch. */
int mccfr_set_max_batch(mccfr_ctx* ctx, uint32_t max_batch);

/* Checkpoint in the reference's .ckpt layout (cfr_gpu.cu:1464-1538):
 * u64 magic 0x43465247505500, u32 table_size, u32 num_actions (row width),
 * f32 regrets[num_actions][table_size], f32 strategy_sum[num_actions][table_size].
 * Loading checks table_size and num_actions against the context
 * (MCCFR_E_FORMAT) and refreshes the strategy snapshot before the next launch. */
int mccfr_save_checkpoint(mccfr_ctx* ctx, const char* path);
int mccfr_load_checkpoint(mccfr_ctx* ctx, const char* path);

/* Launch shape of the walk kernel: workgroup size (64, 128, 256, 512 or
 * 1024 lanes) and traversals per lane (1, 2 or 4; a lane walks its traversals
 * one after another, so a wave covers 64, 128 or 256 traversals; 0: one wave
 * per resident wave slot of the device, each lane walking every (slots x 64)-th
 * traversal, for launches of one tenant; launches of several tenants then use
 * 1).  Defaults 256 and 1.  Results do not depend on the shape. */
int mccfr_set_launch_shape(mccfr_ctx* ctx, uint32_t walk_wg, uint32_t traversals_per_lane);
int mccfr_get_launch_shape(const mccfr_ctx* ctx, uint32_t* walk_wg, uint32_t* traversals_per_lane);

/* Synthetic table state for benchmarks and tests: regret[slot][a] = a uniform
 * (0, 1] draw of stream (seed, id = slot, tag 0x72616E64) for the menu's
 * actions (0 in padding), strategy sums 0, snapshot refreshed.  Paths sampled
 * from it mix folds, calls, raises and all-ins like a partly trained table. */
int mccfr_randomize_tables(mccfr_ctx* ctx, uint64_t seed);

/* Waits for all work of the context. */
int mccfr_sync(mccfr_ctx* ctx);

/* Table geometry: slots, padded actions per row. */
int mccfr_table_shape(const mccfr_ctx* ctx, uint64_t* slots, uint32_t* row_actions);

/* Table transfer.  layout 0 = slot-major rows (AoS, the device layout),
 * 1 = action-major (SoA, the reference's checkpoint layout, cfr_gpu.cu:1479-1487).
 * Sizes are element counts (slots * row_actions). Any pointer may be NULL. */
int mccfr_download_tables(mccfr_ctx* ctx, float* regrets, float* strategy_sum,
                          uint16_t* strategy_bf16, size_t elems, int layout);
int mccfr_upload_tables(mccfr_ctx* ctx, const float* regrets, const float* strategy_sum,
                        size_t elems, int layout);

/* Average-strategy export in the reference's .bin v2 layout (strategy.cpp:6-52):
 * u64 magic 0x53545241540100, u32 table_size, u32 num_actions (the row width),
 * u32 entries, then per entry u32 key (the slot) and f32 probs[num_actions].
 * The rows written are get_strategy's (cfr_gpu.cu:1383-1412): normalised rows
 * whose probabilities add to >= 1e-6, in ascending slot order. */
int mccfr_save_strategy(mccfr_ctx* ctx, uint32_t tenant, const char* path);
/* Writes host rows in the same layout (strategy.cpp:35-52 writes a host table):
 * keys[entries], probs[entries x num_actions]. */
int mccfr_write_strategy(const char* path, uint32_t table_size, uint32_t num_actions, uint32_t entries,
                         const uint32_t* keys, const float* probs);
/* Reads a .bin v2 file (strategy.cpp:54-105).  With keys == NULL only the
 * header is returned and *entries counts the rows a full read returns; else up
 * to `capacity` rows go to keys / probs (probs: entries x num_actions floats).
 * Rows that are uniform 1/num_actions within 1e-6 (placeholders of older
 * exports) are skipped, as the reference's loader does.  A checkpoint or a v1
 * file is rejected with MCCFR_E_FORMAT. */
int mccfr_load_strategy(const char* path, uint32_t* table_size, uint32_t* num_actions, uint32_t* entries,
                        uint32_t* keys, float* probs, size_t capacity);

/* Strategy-quality harness: evaluate_strategy_np (eval.cpp:581-627).  Seat 0
 * plays the normalised average strategy, seats 1..n-1 the scripted archetype
 * (eval.h:5-13), for n_hands hands with the dealer button at hand % n and the
 * game's blinds and starting stack.  Hand h draws from its own Philox stream
 * keyed by `seed` (a fresh 2n+5-card Fisher-Yates deal, then the decisions'
 * uniforms); the reference plays one std::mt19937 sequence instead, so its
 * results are matched in distribution, not hand for hand.  Buckets are the
 * harness's (abstraction.cpp:191-207: clamped at 127), a strategy row counts
 * when it adds to >= 1e-6 (cfr_gpu.cu:1398-1401), else the passive fallback
 * plays.  per_hand (host, n_hands entries, or NULL) gets seat 0's net chips. */
#define MCCFR_OPP_CALLING_STATION 0
#define MCCFR_OPP_NIT 1
#define MCCFR_OPP_MANIAC 2
#define MCCFR_OPP_BALANCED 3
#define MCCFR_OPP_RANDOM 4
typedef struct mccfr_eval_result {
  int64_t hands_played;
  int64_t net_chips;  /* seat 0 over all hands */
  double net_bb;      /* net_chips / big blind */
  double bb_per_100;  /* eval.cpp:620-623 */
} mccfr_eval_result;
int mccfr_evaluate_strategy(mccfr_ctx* ctx, uint32_t tenant, int n_players, int opponent, int64_t n_hands,
                            uint64_t seed, int32_t* per_hand, mccfr_eval_result* out);
/* Same for a strategy given as rows, as mccfr_load_strategy returns them
 * (keys[n_rows], probs[n_rows][num_actions]; num_actions = the game's row
 * width, 8 for the reference game), over the game `cfg`. */
int mccfr_evaluate_rows(mccfr_ctx* ctx, const mccfr_game_cfg* cfg, const uint32_t* keys, const float* probs,
                        uint32_t n_rows, uint32_t num_actions, int n_players, int opponent, int64_t n_hands,
                        uint64_t seed, int32_t* per_hand, mccfr_eval_result* out);

/* Average strategy S/sum(S) per slot, zero rows where sum <= 1e-7
 * (kernel_normalize_strategy, cfr_gpu.cu:1029-1058), into a host array. */
int mccfr_normalized_strategy(mccfr_ctx* ctx, float* out, size_t elems, int layout);

/* Running counters: traversals issued, launches issued, next iteration. */
int mccfr_counters(const mccfr_ctx* ctx, uint64_t* traversals, uint64_t* launches,
                   int64_t* next_iteration);

/* Kernel timing of the last `n` launches issued with timing on (ms, device
 * time between HIP events recorded on the context stream): deal (k_deal),
 * walk (k_walk) and table update (showdown payoffs, record emission, sort and
 * reduce).  Any output pointer may be NULL. */
int mccfr_set_timing(mccfr_ctx* ctx, int enabled);
int mccfr_kernel_times(mccfr_ctx* ctx, float* deal_ms, float* walk_ms, float* update_ms, uint32_t n);

/* Work of the last launch (waits for it): update-player records sorted,
 * runs (table slots) longer than 16 records, and their 256-record pieces. */
int mccfr_last_launch_stats(mccfr_ctx* ctx, uint32_t* records, uint32_t* long_runs, uint32_t* pieces);

/* Device copy of `bytes` bytes between two fresh buffers, `reps` times
 * (calibration of the HBM byte counters; returns ms per copy). */
int mccfr_copy_probe(mccfr_ctx* ctx, size_t bytes, int reps, float* ms_per_copy);

/* `reads` random 16-byte row reads from a fresh table of table_bytes bytes (a
 * power of two), each row's word sum stored as 4 bytes; `reps` passes
 * (calibration of the HBM byte counters for narrow gathers such as the
 * walk's strategy rows; returns ms per pass). */
int mccfr_gather_probe(mccfr_ctx* ctx, size_t table_bytes, size_t reads, int reps, float* ms_per_pass);

/* Writes n hands of 7 distinct uniform cards (rows of 8 bytes, 8th byte 0) to
 * device memory: hand i is the first 7 positions of a Fisher-Yates shuffle
 * driven by draw stream (seed, id = i, tag). */
int mccfr_random_hands_device(mccfr_ctx* ctx, uint64_t seed, uint32_t tag, uint8_t* d_cards8, size_t n);

/* Device-memory plumbing for hosts without their own HIP bindings.
 * kind: 0 host->device, 1 device->host, 2 device->device (synchronous). */
int mccfr_device_alloc(mccfr_ctx* ctx, size_t bytes, void** out);
int mccfr_device_free(mccfr_ctx* ctx, void* p);
int mccfr_memcpy(mccfr_ctx* ctx, void* dst, const void* src, size_t bytes, int kind);

/* ===========================================================================
 * Tenants and batched launches.  No reference counterpart: the reference
 * trains one game per process from the preflop root (cfr_gpu.cu:931-981,
 * 1185-1198).  A context holds several tenants (game configuration, root,
 * tables, update-player schedule) sharing its hand table and launch scratch;
 * one batched launch runs traversals of several tenants together.  Tenant 0 is
 * the context's own game, the one every call above works on.
 * ======================================================================== */

/* Root of a tenant's game tree.
 * street 0: the reference's preflop root (fresh deal of 2N+5 cards, dealer
 *   drawn after it, blinds posted; cfr_gpu.cu:931-981); other fields ignored.
 * street 1/2/3: a public flop/turn/river state (subgame re-solve).  board[0..2],
 *   [0..3] or [0..4] are fixed; every seat's hole cards and the missing board
 *   cards are dealt from the remaining deck (Fisher-Yates, 2N + 5 - fixed
 *   draws, no dealer draw).  The betting round starts with no bets, from the
 *   first seat after `dealer` that is in the hand with chips behind.  Folded
 *   seats (not in active_mask) keep their invested chips in the pot.  Payoffs
 *   count all invested chips: pot = sum(invested) + chips bet after the root.
 *   The tenant's update-player schedule is its active seats in seat order. */
typedef struct mccfr_root_state {
  int32_t street;       /* 0 preflop (reference root), 1 flop, 2 turn, 3 river */
  int32_t dealer;       /* button seat, 0..N-1 */
  uint32_t active_mask; /* seats still in the hand; at least two */
  uint32_t history;     /* action-history bits the info-set keys continue from */
  int32_t stack[MCCFR_MAX_PLAYERS];    /* chips behind at the root */
  int32_t invested[MCCFR_MAX_PLAYERS]; /* chips put into the pot before the root */
  uint8_t board[5];     /* suit-major card codes 0..51, distinct */
  uint8_t reserved[3];
} mccfr_root_state;

/* Device bytes of one tenant's tables (FP32 regrets, FP32 strategy sums, BF16
 * snapshot: 10 bytes per padded action per slot). */
int mccfr_tenant_footprint(const mccfr_game_cfg* cfg, uint64_t* bytes);

/* Creates a tenant with zeroed tables (root NULL = the preflop root).  The
 * context's max_batch bounds the traversals of every launch, over all tenants. */
int mccfr_tenant_create(mccfr_ctx* ctx, const mccfr_game_cfg* cfg, const mccfr_root_state* root,
                        uint32_t* tenant);
int mccfr_tenant_destroy(mccfr_ctx* ctx, uint32_t tenant); /* not tenant 0 */

/* One tenant's share of a batched launch. */
typedef struct mccfr_segment {
  uint32_t tenant;
  uint32_t n;            /* traversals, >= 1 */
  int32_t update_player; /* seat whose regrets this share updates */
  uint32_t reserved;
  int64_t t_weight;      /* strategy-sum weight */
  uint64_t first_tid;    /* traversal id of the share's first traversal */
} mccfr_segment;

/* One batched launch: each segment's traversals, then each tenant's table
 * update, exactly as if every segment ran alone through mccfr_run_launch
 * (a tenant's result does not depend on the other segments).  Segments name
 * distinct tenants with the same padded row width (menus of <= 8 or of 9..16
 * actions); each is padded to a multiple of 256 lanes and the padded total must
 * not exceed max_batch.  trace: HOST array of sum(n) records, segment after
 * segment, or NULL. */
int mccfr_run_segments(mccfr_ctx* ctx, const mccfr_segment* segs, uint32_t n_segs, mccfr_trace* trace);

/* Training loop over several tenants: `launches` batched launches replayed
 * from one captured hipGraph.  In every launch tenant tenants[i] runs counts[i]
 * traversals for the next seat of its own schedule; each tenant advances its
 * own iteration count (t_weight = iteration under linear CFR) and traversal
 * ids, continuing across calls. */
int mccfr_train_tenants(mccfr_ctx* ctx, const uint32_t* tenants, const uint32_t* counts, uint32_t n,
                        int64_t launches);

/* A tenant's counters: traversals issued (= its next traversal id), launches,
 * next iteration and next seat of its schedule. */
int mccfr_tenant_counters(const mccfr_ctx* ctx, uint32_t tenant, uint64_t* traversals, uint64_t* launches,
                          int64_t* next_iteration, int32_t* next_player);

/* Per-tenant forms of the table calls above (tenant 0 = the plain calls). */
int mccfr_tenant_table_shape(const mccfr_ctx* ctx, uint32_t tenant, uint64_t* slots, uint32_t* row_actions);
int mccfr_tenant_download(mccfr_ctx* ctx, uint32_t tenant, float* regrets, float* strategy_sum,
                          uint16_t* strategy_bf16, size_t elems, int layout);
int mccfr_tenant_upload(mccfr_ctx* ctx, uint32_t tenant, const float* regrets, const float* strategy_sum,
                        size_t elems, int layout);
int mccfr_tenant_normalized_strategy(mccfr_ctx* ctx, uint32_t tenant, float* out, size_t elems, int layout);
int mccfr_tenant_randomize(mccfr_ctx* ctx, uint32_t tenant, uint64_t seed);

/* Autotuner: times the deal + walk of the given segment mix (tables are not
 * modified) for every launch shape -- walk workgroup 64/128/256/512/1024 x
 * 1/2/4 traversals per lane -- over `reps` launches each, keeps the fastest as
 * the context's shape and reports it (ms = mean deal + walk time). */
int mccfr_autotune(mccfr_ctx* ctx, const mccfr_segment* segs, uint32_t n_segs, int reps, uint32_t* walk_wg,
                   uint32_t* traversals_per_lane, float* ms);

/* ===========================================================================
 * Re-solve server: continuous batching of subgame re-solves on one context.
 * A request is a game configuration, a root and an iteration budget.  It waits
 * until its tables fit in the server's HBM budget and a running place is
 * free, then runs as a tenant: round r = 1, 2, ... gives every active seat one
 * launch of min(batch, iterations - (r-1)*batch) traversals (one iteration =
 * one traversal per active seat; t_weight = r under linear CFR).  Every server
 * step issues one batched launch holding the next launch of every running
 * request, so requests that arrive while others run join the next launch.
 * No reference counterpart.
 * ======================================================================== */
typedef struct mccfr_server mccfr_server;

/* hbm_budget: device bytes the server may hold for request tables (one pool,
 * allocated here); max_running: requests in one batched launch (<= 4096). */
int mccfr_server_create(mccfr_ctx* ctx, uint64_t hbm_budget, uint32_t max_running, mccfr_server** out);
void mccfr_server_destroy(mccfr_server* srv);

/* Queues a request (root must be a postflop or preflop root of cfg).  batch:
 * traversals per launch, 1..max_batch. */
int mccfr_server_submit(mccfr_server* srv, const mccfr_game_cfg* cfg, const mccfr_root_state* root,
                        int64_t iterations, uint32_t batch, uint64_t* request);

/* One scheduling step: retires launches the device has finished, admits
 * waiting requests that fit (first come, first served), and issues the next
 * batched launch if fewer than two are in flight.  Never waits for the device
 * unless `wait` is non-zero, in which case it first waits for the oldest launch
 * in flight.  Reports requests running and waiting. */
int mccfr_server_step(mccfr_server* srv, int wait, uint32_t* running, uint32_t* waiting);

/* state: 0 waiting, 1 running, 2 done (tables final), 3 released, 4 failed
 * (its launch no longer fits the context's launch scratch after
 * mccfr_set_max_batch shrank it: the step that finds it returns MCCFR_E_STATE
 * naming the request, and the request stops once its launches in flight have
 * finished; release it to free its tables). */
int mccfr_server_poll(mccfr_server* srv, uint64_t request, int* state, uint64_t* traversals);

/* A done request's tenant (for mccfr_tenant_download and the like) and its
 * average strategy (normalized strategy sums, host array). */
int mccfr_server_tenant(mccfr_server* srv, uint64_t request, uint32_t* tenant);
int mccfr_server_result(mccfr_server* srv, uint64_t request, float* strategy, size_t elems, int layout);

/* Frees a done or failed request's tables. */
int mccfr_server_release(mccfr_server* srv, uint64_t request);

#ifdef __cplusplus
}
#endif
#endif /* MCCFR_MI355X_H */
