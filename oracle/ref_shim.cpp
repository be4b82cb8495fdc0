// ref_shim.cpp — C entry points over the reference's own CPU sources.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference's
// poker_cuda/src/hand_eval.cpp and poker_cuda/src/abstraction.cpp, where they
// lie under the mounted reference tree, into oracle/_ref/libref.so (see
// oracle/Makefile, target `ref`).  Used to pin the oracle: the reference's
// table loader/validator and walks run on the generated table, and its
// preflop buckets and action rules produce the golden fixtures.
#include "abstraction.h"
#include "hand_eval.h"

extern "C" {

// hand_eval.cpp:88-102 (size + royal-flush validation, then keeps the table)
int ref_hand_eval_init(const char* path) { return hand_eval_init(path) ? 1 : 0; }

// hand_eval.cpp:134-143
int ref_evaluate_7cards(const unsigned char* c) {
  return evaluate_7cards(c[0], c[1], c[2], c[3], c[4], c[5], c[6]);
}

// hand_eval.cpp:149-173 (n_comm 3 -> eval5, 4 -> eval6, 5 -> eval7)
int ref_evaluate_best(const unsigned char* hole2, const unsigned char* comm, int n_comm) {
  return evaluate_best(hole2[0], hole2[1], comm, n_comm);
}

// abstraction.cpp:42-59
void ref_abstraction_init() { abstraction_init(); }
int ref_preflop_bucket(int h0, int h1) { return preflop_bucket((Card)h0, (Card)h1); }

// abstraction.cpp:167-234
int ref_valid_actions_mask(int pot, int stack, int to_call, int current_bet, int lfr, int bb) {
  return valid_actions_mask(pot, stack, to_call, current_bet, lfr, bb);
}
int ref_action_to_chips(int a, int pot, int stack, int to_call, int current_bet, int lfr, int bb) {
  return action_to_chips((Action)a, pot, stack, to_call, current_bet, lfr, bb);
}

// abstraction.cpp:191-207 (the strategy harness's buckets: Chen percentile
// preflop, clamped rank buckets on flop / turn / river)
void ref_precompute_buckets(const unsigned char* hole, const unsigned char* community, int n, int* hb, int* bb) {
  precompute_buckets(hole, community, n, hb, bb);
}
}
