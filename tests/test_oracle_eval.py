"""The oracle's strategy-quality harness (eval.cpp:45-627), pinned on the CPU:

* its buckets equal the reference's own precompute_buckets (golden fixture
  written from abstraction.cpp / hand_eval.cpp, tests/golden/eval_buckets.npz);
* known-answer hands with given cards and decisions, worked out by hand from
  the rules of no-limit hold'em (blinds, action order, reopening, pot-fraction
  raises, all-in runout, split pots with an odd chip, folded money in a side pot);
* chip conservation over many random hands (no chips created; a split loses
  at most its odd chips);
* the archetype action frequencies documented at eval.h:5-13;
* the reference's own random stream (std::mt19937, persistent deck) and the
  project's per-hand Philox stream play the same game: equal mean results
  within sampling error.
"""
import os

import numpy as np
import pytest

import mccfr
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FOLD, CHECK, CALL, RMIN, RHALF, R23, RPOT, ALLIN = range(8)


def card(s):
    return "cdhs".index(s[1]) * 13 + "23456789TJQKA".index(s[0])


def deal(holes, board):
    return np.array([card(c) for h in holes for c in h.split()] + [card(c) for c in board.split()], np.uint8)


@pytest.fixture(scope="module")
def hr(hand_table):
    return hand_table


def test_buckets_match_the_reference_precompute_buckets(hr, golden_lut):
    g = np.load(os.path.join(GOLDEN, "eval_buckets.npz"))
    for i in range(len(g["cards"])):
        hb, bk = oracle.eval_buckets(hr, golden_lut, 9, g["cards"][i])
        assert np.array_equal(hb, g["hole_buckets"][i])
        assert np.array_equal(bk, g["board_buckets"][i])
    # the harness clamps rank-0 walks to 127 where the training walk truncates
    assert (g["board_buckets"][:, :, 1:3] == 127).any()


def scripted(hr, lut, n, dealer, cards, actions, sb=10, bb=20, stack=1000):
    cfg = mccfr.reference_cfg(num_players=n, starting_stack=stack, small_blind=sb, big_blind=bb)
    return oracle.eval_scripted(cfg, hr, lut, n, dealer, cards, actions)


BOARD_DRY = "2c 7d 9h Js Kd"
BOARD_BROADWAY = "Tc Jd Qh Ks Ac"


@pytest.mark.parametrize("name,n,dealer,holes,board,actions,sb,bb,used,net", [
    # heads-up: the dealer posts the small blind and acts first preflop
    ("sb_folds", 2, 0, ["As Ah", "Qc Qh"], BOARD_DRY, [FOLD], 10, 20, 1, [-10, 10]),
    # three-way limp, checked down; aces win a 60-chip pot
    ("limp_checkdown", 3, 0, ["As Ah", "Qc Qh", "3c 4d"], BOARD_DRY,
     [CALL, CALL, CHECK] + [CHECK] * 9, 10, 20, 12, [40, -20, -20]),
    # both play the board straight: 40 chips split
    ("board_chop", 2, 0, ["2c 3d", "2h 3s"], BOARD_BROADWAY, [CALL, CHECK] + [CHECK] * 6, 10, 20, 8, [0, 0]),
    # blinds 5/10; the small blind folds its 5 into the pot; the tie splits
    # 15 (7 + 7, one chip lost) and then 10 (5 + 5)
    ("odd_chip", 3, 0, ["2c 3d", "4c 5d", "2h 3s"], BOARD_BROADWAY, [CALL, FOLD, CHECK] + [CHECK] * 6, 5, 10, 9,
     [2, -5, 2]),
    # min-raise to 40, min-reraise to 60 reopens the action, call; 120 to the aces
    ("reopen", 2, 0, ["3c 4d", "As Ah"], BOARD_DRY, [RMIN, RMIN, CALL] + [CHECK] * 6, 10, 20, 9, [-60, 60]),
    # pot raise to 50 (pot 30), call; flop half-pot bet 50 into 100, two-thirds
    # raise to 150 (pot 150), fold
    ("pot_fractions", 2, 0, ["As Ah", "3c 4d"], BOARD_DRY, [RPOT, CALL, RHALF, R23, FOLD], 10, 20, 5, [100, -100]),
    # pot raise to 50 is a 30-chip raise, so the min-reraise goes to 50 + 30 = 80
    ("lfr_tracks_raise", 2, 0, ["3c 4d", "As Ah"], BOARD_DRY, [RPOT, RMIN, CALL] + [CHECK] * 6, 10, 20, 9,
     [-80, 80]),
    # all-in and call preflop: no more decisions, the board runs out
    ("allin_runout", 2, 0, ["As Ah", "Qc Qh"], BOARD_DRY, [ALLIN, ALLIN], 10, 20, 2, [1000, -1000]),
    # six-max, dealer 3: blinds on seats 4 and 5, seat 0 first; all fold to the big blind
    ("fold_to_bb", 6, 3, ["2c 3d", "4c 5d", "6c 7d", "8c 9d", "Tc Jd", "Qc Kd"], "2h 3h 4h 5s 9s",
     [FOLD] * 5, 10, 20, 5, [0, 0, 0, 0, -10, 10]),
])
def test_known_answer_hands(hr, golden_lut, name, n, dealer, holes, board, actions, sb, bb, used, net):
    got_used, got_net = scripted(hr, golden_lut, n, dealer, deal(holes, board), actions, sb=sb, bb=bb)
    assert got_used == used, name
    assert got_net.tolist() == net, name


def test_illegal_decision_is_reported(hr, golden_lut):
    # facing the big blind, the small blind cannot check
    used, _ = scripted(hr, golden_lut, 2, 0, deal(["As Ah", "Qc Qh"], BOARD_DRY), [CHECK])
    assert used == -2
    # decisions run out
    used, _ = scripted(hr, golden_lut, 2, 0, deal(["As Ah", "Qc Qh"], BOARD_DRY), [CALL])
    assert used == -2


def random_strategy(log2, row=8, seed=0, density=0.9):
    rng = np.random.default_rng(seed)
    P = rng.random((1 << log2, row)).astype(np.float32)
    P[rng.random(1 << log2) > density] = 0.0
    P /= np.maximum(P.sum(1, keepdims=True), 1e-30)
    return P.astype(np.float32)


@pytest.mark.parametrize("n", [2, 3, 6, 9])
@pytest.mark.parametrize("opp", range(5))
def test_chip_conservation(hr, golden_lut, n, opp):
    cfg = mccfr.reference_cfg(num_players=min(n, 6), table_log2=12)
    P = random_strategy(12, seed=n * 7 + opp)
    bot, seats = oracle.eval_hands(cfg, hr, golden_lut, P, n, opp, 3000, seed=11, seats=True, threads=4)
    total = seats.sum(1)
    assert (total <= 0).all()
    assert (total >= -n * n).all()
    assert (seats >= -1000).all()
    assert np.array_equal(bot, seats[:, 0])
    assert np.count_nonzero(total) < len(total) // 10


def test_archetype_frequencies(golden_lut):
    """eval.h:5-13: calling station 8/72/20, nit 60/30/10, maniac 10/20/70,
    balanced 28/44/28 (fold / call / raise, raises split evenly; a free check
    takes the fold and call shares); random: uniform over legal actions."""
    cfg = mccfr.reference_cfg()
    draws = 400_000
    facing = (1 << FOLD) | (1 << CALL) | (1 << RMIN) | (1 << RHALF) | (1 << RPOT) | (1 << ALLIN)
    free = (1 << CHECK) | (1 << RMIN) | (1 << ALLIN)
    no_raise = (1 << FOLD) | (1 << CALL)
    for opp, (f, c, r) in enumerate([(0.08, 0.72, 0.20), (0.60, 0.30, 0.10), (0.10, 0.20, 0.70),
                                     (0.28, 0.44, 0.28)]):
        for vm, to_call, expect in [
            (facing, 20, {FOLD: f, CALL: c, RMIN: r / 4, RHALF: r / 4, RPOT: r / 4, ALLIN: r / 4}),
            (free, 0, {CHECK: f + c, RMIN: r / 2, ALLIN: r / 2}),
            (no_raise, 20, {FOLD: f, CALL: c + r}),
        ]:
            counts = oracle.eval_scripted_counts(cfg, opp, vm, to_call, draws, seed=opp + 1)
            assert counts.sum() == draws
            for a in range(16):
                p = expect.get(a, 0.0)
                sd = np.sqrt(draws * p * (1 - p)) + 1
                assert abs(counts[a] - draws * p) < 5 * sd, (opp, vm, a, counts[a], draws * p)
    counts = oracle.eval_scripted_counts(cfg, oracle.RANDOM, facing, 20, draws)
    sel = [FOLD, CALL, RMIN, RHALF, RPOT, ALLIN]
    assert counts[sel].sum() == draws
    assert (np.abs(counts[sel] - draws / 6) < 5 * np.sqrt(draws / 6)).all()


def test_reference_stream_and_philox_stream_agree_in_distribution(hr, golden_lut):
    cfg = mccfr.reference_cfg(num_players=2, table_log2=10)
    P = random_strategy(10, seed=3)
    n = 200_000
    a = oracle.eval_hands(cfg, hr, golden_lut, P, 2, oracle.RANDOM, n, seed=42, stream=1).astype(np.float64)
    b = oracle.eval_hands(cfg, hr, golden_lut, P, 2, oracle.RANDOM, n, seed=42, stream=0, threads=8)
    b = b.astype(np.float64)
    se = np.sqrt(a.var() / n + b.var() / n)
    assert abs(a.mean() - b.mean()) < 5 * se, (a.mean(), b.mean(), se)
    # the reference stream is one sequence: the same seed replays it
    a2 = oracle.eval_hands(cfg, hr, golden_lut, P, 2, oracle.RANDOM, 1000, seed=42, stream=1)
    assert np.array_equal(a2, a[:1000].astype(np.int32))


def test_philox_hands_are_independent_of_the_split(hr, golden_lut):
    cfg = mccfr.reference_cfg(num_players=6, table_log2=12)
    P = random_strategy(12, seed=5)
    whole = oracle.eval_hands(cfg, hr, golden_lut, P, 6, oracle.BALANCED, 4000, seed=9, threads=1)
    part = oracle.eval_hands(cfg, hr, golden_lut, P, 6, oracle.BALANCED, 1000, seed=9, first_hand=3000, threads=3)
    assert np.array_equal(whole[3000:], part)


def test_empty_strategy_plays_the_passive_fallback(hr, golden_lut):
    """With no rows the strategy seat checks or calls (never folds a free check,
    never raises): its losses are bounded by what it calls."""
    cfg = mccfr.reference_cfg(num_players=2, table_log2=8)
    P = np.zeros((1 << 8, 8), np.float32)
    bot, seats = oracle.eval_hands(cfg, hr, golden_lut, P, 2, oracle.NIT, 5000, seed=1, seats=True)
    assert (seats.sum(1) <= 0).all()
    # a nit folds 60% when facing a bet but the passive seat never bets, so the
    # nit is only ever facing the big blind preflop
    assert np.count_nonzero(bot) > 0
