"""GPU 7-card evaluators (rank-walk table in HBM and arithmetic perfect hash)
against the oracle: bit-exact ranks."""
import os

import numpy as np
import pytest

import mccfr
import oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HAND_TAG = 0x68616E64  # stream family of the hand microbenchmark


@pytest.fixture(scope="module")
def ctx():
    c = mccfr.Context(mccfr.reference_cfg(table_log2=10), 1024)
    c.use_generated_hand_table()
    yield c
    c.close()


@pytest.mark.parametrize("evaluator", [0, 1])
def test_eval7_golden(ctx, evaluator):
    g = np.load(os.path.join(GOLDEN, "hand_walks.npz"))
    ranks = ctx.eval7(g["cards"], evaluator)
    assert np.array_equal(ranks, g["ranks"][:, 3])
    padded = np.zeros((len(g["cards"]), 8), np.uint8)
    padded[:, :7] = g["cards"]
    assert np.array_equal(ctx.eval7(padded, evaluator), g["ranks"][:, 3])


def test_random_hands_match_oracle_deal(ctx):
    n = 4096
    d = ctx.dalloc(n * 8)
    try:
        ctx.random_hands_device(1, HAND_TAG, d, n)
        got = np.empty((n, 8), np.uint8)
        ctx.d2h(got, d)
    finally:
        ctx.dfree(d)
    import ctypes as C
    cards = np.zeros(7, np.uint8)
    dealer = C.c_int()
    for i in range(0, n, 7):
        oracle.lib().orc_deal(1, i, HAND_TAG, 1, oracle.p(cards), C.byref(dealer))
        assert got[i, :7].tolist() == cards.tolist()
    assert (got[:, 7] == 0).all()
    s = np.sort(got[:, :7], axis=1)
    assert (np.diff(s.astype(int), axis=1) > 0).all()


def test_eval7_2pow24_random_hands(ctx, hand_table):
    """2^24 uniform 7-card hands (seed 1): both GPU evaluators equal the oracle walk."""
    n = 1 << 24
    dc = ctx.dalloc(n * 8)
    dr = ctx.dalloc(n * 2)
    try:
        ctx.random_hands_device(1, HAND_TAG, dc, n)
        cards = np.empty((n, 8), np.uint8)
        ctx.d2h(cards, dc)
        want = np.empty(n, np.uint16)
        oracle.lib().orc_walk7_batch(oracle.p(hand_table), oracle.p(cards), 8, oracle.p(want), n, 0)
        for ev in (0, 1):
            ctx.eval7_device(dc, 8, dr, n, ev)
            got = np.empty(n, np.uint16)
            ctx.d2h(got, dr)
            assert np.array_equal(got, want), ev
    finally:
        ctx.dfree(dc)
        ctx.dfree(dr)
