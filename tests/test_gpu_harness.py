"""The strategy-quality harness on the device (evaluate_strategy_np,
eval.cpp:581-627) against the oracle's restatement, hand for hand.

Both run the project's per-hand Philox stream, so seat 0's net chips must be
equal for every hand (integers: bit-exact), over trained tables, dense random
rows, an empty strategy and a custom menu with a raise cap, for 2 to 9 seats
and every archetype.  The oracle itself is pinned on the CPU
(tests/test_oracle_eval.py).
"""
import numpy as np
import pytest

import mccfr
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(hand_table):
    cfg = mccfr.reference_cfg(num_players=3, starting_stack=1000, table_log2=14, seed=3)
    ctx = mccfr.Context(cfg, 8192)
    ctx.set_hand_table(hand_table)
    ctx.train(20, 8192)
    yield ctx
    ctx.close()


def oracle_hands(cfg, hr, lut, P, n, opp, hands, seed, first=0):
    return oracle.eval_hands(cfg, hr, lut, P.reshape(-1, 8 if cfg.num_actions <= 8 else 16), n, opp, hands,
                             seed=seed, first_hand=first, threads=8)


@pytest.mark.parametrize("n", [2, 3, 6, 9])
@pytest.mark.parametrize("opp", range(5))
def test_trained_strategy_hand_for_hand(trained, hand_table, golden_lut, n, opp):
    hands, seed = 6000, 100 + 10 * n + opp
    print(f"harness n={n} opp={opp} hands={hands}", flush=True)
    res, got = trained.evaluate_strategy(n, opp, hands, seed=seed, per_hand=True)
    P = trained.normalized_strategy()
    want = oracle_hands(trained.cfg, hand_table, golden_lut, P, n, opp, hands, seed)
    assert np.array_equal(got, want), np.nonzero(got != want)[0][:10]
    assert res.hands_played == hands and res.net_chips == int(want.sum())
    assert res.net_bb == pytest.approx(res.net_chips / 20)
    assert res.bb_per_100 == pytest.approx(res.net_chips / 20 / hands * 100)
    # the strategy seat reaches its rows: not every decision is the fallback
    assert np.count_nonzero(want) > hands // 4


def dense_rows(log2, row, seed, density=0.8):
    rng = np.random.default_rng(seed)
    keep = np.nonzero(rng.random(1 << log2) < density)[0].astype(np.uint32)
    probs = rng.random((keep.size, row)).astype(np.float32)
    probs /= probs.sum(1, keepdims=True)
    P = np.zeros((1 << log2, row), np.float32)
    P[keep] = probs
    return keep, probs.astype(np.float32), P


@pytest.mark.parametrize("case", ["reference", "menu_cap", "empty"])
def test_rows_hand_for_hand(trained, hand_table, golden_lut, case):
    if case == "menu_cap":
        menu = [(mccfr.FOLD, 0, 1), (mccfr.CHECK_CALL, 0, 1), (mccfr.POT_FRAC, 1, 3), (mccfr.POT_FRAC, 1, 1),
                (mccfr.POT_FRAC, 5, 2), (mccfr.ALL_IN, 0, 1)]
        cfg = mccfr.menu_cfg(6, 2000, menu, max_raises=2, table_log2=12, seed=0)
    else:
        cfg = mccfr.reference_cfg(num_players=6, table_log2=12)
    row = 8
    if case == "empty":
        keys, probs, P = np.zeros(0, np.uint32), np.zeros((0, row), np.float32), np.zeros((1 << 12, row), np.float32)
    else:
        keys, probs, P = dense_rows(12, row, seed=7)
        probs[:, cfg.num_actions:] = 0
        P[:, cfg.num_actions:] = 0
    for n, opp in [(2, mccfr.RANDOM), (6, mccfr.BALANCED), (9, mccfr.MANIAC), (4, mccfr.NIT)]:
        print(f"harness rows {case} n={n} opp={opp}", flush=True)
        res, got = trained.evaluate_rows(cfg, keys, probs, n, opp, 5000, seed=9, per_hand=True)
        want = oracle_hands(cfg, hand_table, golden_lut, P, n, opp, 5000, 9)
        assert np.array_equal(got, want), (case, n, opp, np.nonzero(got != want)[0][:10])
        assert res.net_chips == int(want.sum())


def test_launch_chunks_and_totals(trained, hand_table, golden_lut):
    """More hands than one launch holds (2^22): the second launch continues the
    hand numbering; the device total equals the per-hand sum."""
    hands = (1 << 22) + 3000
    print(f"harness chunks hands={hands}", flush=True)
    res, got = trained.evaluate_strategy(6, mccfr.CALLING_STATION, hands, seed=77, per_hand=True)
    assert res.net_chips == int(got.astype(np.int64).sum())
    P = trained.normalized_strategy()
    for first in (0, (1 << 22) - 1000, (1 << 22)):
        want = oracle_hands(trained.cfg, hand_table, golden_lut, P, 6, mccfr.CALLING_STATION, 2000, 77, first)
        assert np.array_equal(got[first:first + 2000], want), first
    res2 = trained.evaluate_strategy(6, mccfr.CALLING_STATION, hands, seed=77)
    assert res2.net_chips == res.net_chips


def test_bad_arguments(trained):
    for args in [(1, 0, 10), (10, 0, 10), (3, 5, 10), (3, 0, -1)]:
        with pytest.raises(mccfr.MccfrError) as ei:
            trained.evaluate_strategy(*args)
        assert ei.value.code == mccfr.E_ARG
    res = trained.evaluate_strategy(3, 0, 0)
    assert res.hands_played == 0 and res.net_chips == 0
