/*
 * oracle.h — CPU restatement of the reference MCCFR hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * as the timed CPU baseline.  The product (libmccfr) never links or calls it.
 *
 * Every function restates the reference file:line given beside it.  Parity
 * pins (what checks each function against something other than itself) are
 * listed in oracle.cpp's header.
 */
#ifndef MCCFR_ORACLE_H
#define MCCFR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/mccfr/mccfr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- cards and the rank-walk table (hand_eval.cpp:23-57, 107-143; cfr_gpu.cu:87-150) */
int orc_to_rm(int card);
uint16_t orc_decode_packed(uint16_t packed);
uint16_t orc_walk5(const int32_t* hr, const uint8_t* c);
uint16_t orc_walk6(const int32_t* hr, const uint8_t* c);
uint16_t orc_walk7(const int32_t* hr, const uint8_t* c);
void orc_walk7_batch(const int32_t* hr, const uint8_t* cards, size_t stride, uint16_t* out,
                     size_t n, int threads);
int orc_validate_table(const int32_t* hr, size_t words); /* hand_eval.cpp:59-86: 1 = ok */

/* ---- independent brute-force evaluator (best 5 of n by enumeration) */
int orc_bf_class_count(int per_category[9]); /* distinct 5-card classes (7462) per category */
int orc_bf_class(const uint8_t* cards, int n);
/* All 133,784,560 seven-card hands: category histogram of the table walk
 * (index 1..9 by decoded class) and, if check_bf, the number of hands whose
 * walk rank differs from the brute-force class. */
int orc_exhaustive7(const int32_t* hr, int threads, int check_bf, uint64_t cat_counts[10],
                    uint64_t* mismatches);
/* All 2,598,960 five-card hands: category histogram of the brute-force class. */
int orc_exhaustive5_bf(uint64_t cat_counts[10]);

/* ---- random streams: Philox4x32-10 (Salmon et al., SC'11) */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t orc_draw_u32(uint64_t seed, uint64_t tid, uint32_t tag, uint32_t k);
float orc_u01(uint32_t x);
uint16_t orc_f32_to_bf16(float x);
float orc_bf16_to_f32(uint16_t h);

/* ---- abstraction (abstraction.cpp:10-59, 144-234; cfr_gpu.cu:165-244, 476-489) */
void orc_preflop_lut(uint8_t out[2704]);
uint32_t orc_info_hash(uint8_t player, uint8_t hole_b, uint8_t board_b, uint8_t street,
                       uint32_t action_bits, int table_log2);
uint32_t orc_valid_mask(const mccfr_game_cfg* cfg, int pot, int stack, int to_call,
                        int current_bet, int last_full_raise, int raises);
int orc_chips(const mccfr_game_cfg* cfg, int a, int pot, int stack, int to_call,
              int current_bet, int last_full_raise);
int orc_passive_fallback(const mccfr_game_cfg* cfg, uint32_t mask);
int orc_board_bucket(const int32_t* hr, const uint8_t* hole2, const uint8_t* comm5, int street);

/* Deal of one traversal (cfr_gpu.cu:460-470, 931-937, 963): 2N+5 cards then dealer. */
void orc_deal(uint64_t seed, uint64_t tid, uint32_t tag, int num_players, uint8_t* cards_out,
              int* dealer_out);

/* ---- one launch replayed in FP64 (cfr_gpu.cu:318-351, 878-1011, 1018-1022).
 * R, S: pre-launch regrets/strategy sums, slot-major rows of `row` floats
 * (row = 8 for menus of <= 8 actions, else 16).  sigma: BF16 snapshot (same
 * shape) or NULL to derive it from R by regret matching.  dR/dS receive the
 * FP64 sum of the launch's contributions, aR/aS the FP64 sum of their absolute
 * values (all four zero-filled by the caller).  trace: n records or NULL.
 * With dR..aS all NULL only the trace is produced (trace must then be given). */
int orc_replay_launch(const mccfr_game_cfg* cfg, const int32_t* hr, const uint8_t* lut,
                      const float* R, const uint16_t* sigma, int update_player,
                      int64_t t_weight, uint64_t first_tid, uint32_t n, uint32_t tag,
                      double* dR, double* dS, double* aR, double* aS, mccfr_trace* trace);
/* Same from a tenant root (mccfr.h mccfr_root_state; NULL or street 0 = the
 * reference's preflop root). */
int orc_replay_launch_root(const mccfr_game_cfg* cfg, const mccfr_root_state* root, const int32_t* hr,
                           const uint8_t* lut, const float* R, const uint16_t* sigma, int update_player,
                           int64_t t_weight, uint64_t first_tid, uint32_t n, uint32_t tag, double* dR,
                           double* dS, double* aR, double* aS, mccfr_trace* trace);
/* Hole cards (2N) and full board (5) of traversal tid from a root. */
void orc_deal_root(const mccfr_game_cfg* cfg, const mccfr_root_state* root, uint64_t tid, uint32_t tag,
                   uint8_t* hole_out, uint8_t* board_out);
/* Sparse form of a launch's contributions: the touched table entries
 * (idx = slot * row + action, ascending) and, per entry, the FP64 sums of the
 * regret and strategy-sum contributions and of their absolute values. */
typedef struct orc_sparse {
  int64_t n;
  uint64_t* idx;
  double *dR, *dS, *aR, *aS;
} orc_sparse;
/* orc_replay_launch_root for full-size tables: traversals split over
 * `threads` host threads (0 = all), contributions returned sparse (free with
 * orc_sparse_free); sigma must be given.  Returns the entry count or -1. */
int64_t orc_replay_sparse(const mccfr_game_cfg* cfg, const mccfr_root_state* root, const int32_t* hr,
                          const uint8_t* lut, const float* R, const uint16_t* sigma, int update_player,
                          int64_t t_weight, uint64_t first_tid, uint32_t n, uint32_t tag, int threads,
                          mccfr_trace* trace, orc_sparse* out);
void orc_sparse_free(orc_sparse* s);

/* One decision node of a walk, as the walk saw it and left it (the state
 * fields are those after the action was applied and the queue rebuilt). */
typedef struct orc_decision {
  uint8_t player;
  uint8_t street;
  uint8_t action;
  uint8_t is_update;    /* 1 when player == the update player */
  uint32_t legal;       /* legal-action mask at the node */
  uint32_t history;     /* action-history bits before the action */
  uint32_t slot;        /* info-set slot */
  int32_t chips;        /* chips the action put in */
  int32_t pot;          /* pot after the action */
  int32_t current_bet;  /* street's bet to match after the action */
  int32_t lfr;          /* last full raise after the action */
  int32_t raises;       /* aggressive actions this street after the action */
  int32_t n_queue;      /* seats left to act after the action ... */
  uint32_t queue;       /* ... in order, 4 bits each from the low bits */
} orc_decision;

/* One traversal with given cards, dealer and tables (the walk of
 * orc_replay_launch_root after its deal: cfr_gpu.cu:954-981 and 644-860).
 * hole: 2N cards seat by seat; board: 5 cards (a postflop root's fixed cards
 * must be its first street + 2); dealer: the preflop button (ignored for a
 * postflop root, which names its own).  The sampling draws come from stream
 * (cfg->seed, tid, tag) after the deal's draws, as in a replayed launch.
 * trace: one record or NULL; log: up to log_cap decisions, *n_log their
 * number; contributions (slot, action, regret delta, strategy-sum delta) of
 * the update player's nodes: up to c_cap entries, *n_contrib their number.
 * Returns 0, or -1 on a bad argument / a capacity exceeded. */
int orc_walk_cards(const mccfr_game_cfg* cfg, const mccfr_root_state* root, const int32_t* hr, const uint8_t* lut,
                   const uint8_t* hole, const uint8_t* board, int dealer, const float* R, const uint16_t* sigma,
                   int update_player, int64_t t_weight, uint64_t tid, uint32_t tag, mccfr_trace* trace,
                   orc_decision* log, int log_cap, int* n_log, uint32_t* c_slot, uint32_t* c_action, float* c_rd,
                   float* c_sd, int c_cap, int* n_contrib);

/* Regret matching of slot-major rows (kernel_regret_matching, cfr_gpu.cu:318-351). */
void orc_regret_matching(const float* R, uint16_t* sigma, uint64_t slots, int row);

/* ---- the timed CPU baseline: the reference launch sequence in FP32 with
 * action-major (SoA) tables: full regret-matching sweep, traversals split over
 * `threads` threads with private delta lists, merge, full CFR+ clamp sweep.
 * Runs `launches` launches of `batch` traversals cycling the update player
 * from `first_player`; returns seconds spent (steady clock), or < 0 on error. */
double orc_cpu_baseline(const mccfr_game_cfg* cfg, const int32_t* hr, const uint8_t* lut,
                        float* R_soa, float* S_soa, uint16_t* sigma_soa, int launches,
                        uint32_t batch, int first_player, int64_t first_iter,
                        uint64_t first_tid, int threads);

/* ---- strategy-quality harness (eval.cpp:45-627).  Seat 0 plays the
 * normalised strategy P (slot-major rows of `row`; a row is present when it
 * adds to >= 1e-6, as get_strategy keeps it), seats 1..n-1 the archetype
 * `opponent` (0 calling station, 1 nit, 2 maniac, 3 balanced, 4 random);
 * dealer = hand % n.  stream 0: the project's per-hand Philox stream, hands
 * first_hand.. (threads > 1 splits them); stream 1: the reference's single
 * std::mt19937(seed) stream with a persistent deck (first_hand must be 0).
 * bot_net[n_hands] / seat_net[n_hands][n] (either may be NULL) get net chips. */
int orc_eval_hands(const mccfr_game_cfg* cfg, const int32_t* hr, const uint8_t* lut, const float* P, int row,
                   int n_players, int opponent, int stack, uint64_t seed, uint64_t first_hand, int64_t n_hands,
                   int stream, int threads, int32_t* bot_net, int32_t* seat_net);
/* One hand with the cards (2n hole cards seat by seat, then 5 board cards) and
 * every decision given in order; returns decisions used, -2 if an action was
 * illegal or the list ran out. */
int orc_eval_scripted(const mccfr_game_cfg* cfg, const int32_t* hr, const uint8_t* lut, int n_players, int stack,
                      int dealer, const uint8_t* cards, const int32_t* actions, int n_actions, int32_t* seat_net);
/* Counts of `draws` archetype decisions at legal mask vm (draw i uses hand stream i). */
int orc_eval_scripted_action(const mccfr_game_cfg* cfg, int opponent, uint32_t vm, int to_call, uint64_t seed,
                             int64_t draws, int64_t* counts);
/* The harness's buckets (abstraction.cpp:191-207): hb[n], bk[n][4]. */
int orc_eval_buckets(const int32_t* hr, const uint8_t* lut, int n_players, const uint8_t* cards, int32_t* hb,
                     int32_t* bk);

#ifdef __cplusplus
}
#endif
#endif
