"""GPU traversal launches against the oracle's FP64 replay, launch by launch
(teacher-forced from the GPU's own pre-launch tables):

* per traversal, bit-exact: terminal utility (FP32 bits), path hash over every
  (slot, action) decision, decision and update-node counts, terminal kind;
* the BF16 strategy snapshot equals regret matching of the FP32 regrets, bit-exact;
* regrets and strategy sums after the launch satisfy
  |x - x_hat| <= 1e-5 * max(|x_hat|, sum|delta|), x_hat = pre + FP64 sum of contributions.
"""
import numpy as np
import pytest

import helpers
import mccfr
import oracle

pytestmark = pytest.mark.gpu


def run_stepwise(cfg, hand_table, lut, n, launches, first_tid=0, random_seed=None):
    row = 8 if cfg.num_actions <= 8 else 16
    N = cfg.num_players
    with mccfr.Context(cfg, n) as ctx:
        ctx.set_hand_table(hand_table)
        if random_seed is not None:
            ctx.randomize_tables(random_seed)
        tid = first_tid
        for L in range(launches):
            up, t = L % N, L // N + 1
            w = t if cfg.use_linear_cfr else 1
            R0, S0, sig0 = ctx.download()
            assert np.array_equal(sig0, oracle.regret_matching(R0, row)), f"snapshot, launch {L}"
            print(f"launch {L}: up={up} t={w} tid={tid} n={n}", flush=True)
            tr = ctx.run_launch(up, w, tid, n, trace=True)
            dR, dS, aR, aS, otr = oracle.replay_launch(cfg, hand_table, lut, R0, sig0, up, w, tid, n)
            for f in ("path_hash", "decisions", "updates", "terminal"):
                assert np.array_equal(tr[f], otr[f]), (f, L)
            assert np.array_equal(tr["utility"].view(np.uint32), otr["utility"].view(np.uint32)), L
            R1, S1, sig1 = ctx.download()
            eR, eS = helpers.expected_post(R0, S0, dR, dS, bool(cfg.use_cfr_plus))
            okR = helpers.within_tol(R1, eR, aR)
            okS = helpers.within_tol(S1, eS, aS)
            assert okR.all(), (L, np.argwhere(~okR)[:5])
            assert okS.all(), (L, np.argwhere(~okS)[:5])
            tid += n
        R, S, sig = ctx.download()
        assert np.array_equal(sig, oracle.regret_matching(R, row))
        return R, S, tid


CONFIGS = {
    "ref_2p_50bb": lambda: mccfr.reference_cfg(num_players=2, starting_stack=1000, table_log2=14, seed=1),
    "ref_3p_20bb": lambda: mccfr.reference_cfg(num_players=3, starting_stack=400, table_log2=14, seed=2),
    "ref_6p_100bb": lambda: mccfr.reference_cfg(num_players=6, starting_stack=2000, table_log2=16, seed=3),
    "3p_20bb_fcpa_cap2": lambda: mccfr.cfg_3p_20bb(table_log2=14, seed=4),
    "6p_100bb_menu6_cap3": lambda: mccfr.cfg_6p_100bb(table_log2=16, seed=5),
    "6p_200bb_menu10_cap4": lambda: mccfr.cfg_6p_200bb_10(table_log2=16, seed=6),
    "ref_2p_short_allin": lambda: mccfr.reference_cfg(num_players=2, starting_stack=30, table_log2=12, seed=7),
    "ref_4p_no_cfrplus_no_lcfr": lambda: mccfr.reference_cfg(num_players=4, starting_stack=1000,
                                                             use_cfr_plus=False, use_linear_cfr=False,
                                                             table_log2=14, seed=8),
}


@pytest.mark.parametrize("start", ["zero", "random"])
@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_stepwise_parity(name, start, hand_table, golden_lut):
    cfg = CONFIGS[name]()
    R, S, _ = run_stepwise(cfg, hand_table, golden_lut, n=4096, launches=3 * cfg.num_players,
                           random_seed=None if start == "zero" else 99)
    assert np.count_nonzero(S) > 0
    if cfg.use_cfr_plus:
        assert (R >= 0).all()


EDGE_CASES = {
    # one traversal: a launch of 256 lanes with 255 padding lanes
    "one_traversal": (lambda: mccfr.reference_cfg(num_players=3, starting_stack=400, table_log2=14, seed=31), 1, 3),
    # a ragged count: a partial last wave and emission block
    "ragged_257": (lambda: mccfr.cfg_6p_100bb(table_log2=16, seed=32), 257, 6),
    # 16-slot tables: ~6,000 records per slot, so runs span several 4,096-record
    # region groups, one 8-bit radix pass sorts the 4-bit keys, and the last run
    # ends at the last record
    "sixteen_slots": (lambda: mccfr.cfg_6p_100bb(table_log2=4, seed=33), 1 << 16, 6),
}


@pytest.mark.parametrize("case", sorted(EDGE_CASES))
def test_edge_shapes_parity(case, hand_table, golden_lut):
    make, n, launches = EDGE_CASES[case]
    cfg = make()
    R, S, _ = run_stepwise(cfg, hand_table, golden_lut, n=n, launches=launches, random_seed=7)
    assert np.count_nonzero(S) > 0


def test_min_raise_war_fills_emission_blocks_beyond_their_stage(hand_table, golden_lut):
    """Heads-up, every slot's regret row one-hot on the min-raise: both seats
    min-raise until the depth guard, so every traversal holds 10 update-player
    nodes and each 256-lane emission block 2,560 records, more than the
    emission's LDS stage (k_emit writes them directly).  Checked against the
    oracle like every other launch."""
    cfg = mccfr.menu_cfg(2, 4000, [(mccfr.FOLD, 0, 1), (mccfr.CHECK_CALL, 0, 1), (mccfr.RAISE_MIN, 0, 1),
                                   (mccfr.ALL_IN, 0, 1)], max_raises=0, table_log2=12, seed=34)
    n, row = 2048, 8
    with mccfr.Context(cfg, n) as ctx:
        ctx.set_hand_table(hand_table)
        R = np.zeros((1 << cfg.table_log2, row), np.float32)
        R[:, 2] = 1.0
        ctx.upload(R.reshape(-1), np.zeros(R.size, np.float32))
        for L in range(2):
            R0, S0, sig0 = ctx.download()
            tr = ctx.run_launch(L, 1, L * n, n, trace=True)
            assert (tr["updates"] == 10).all(), np.bincount(tr["updates"])
            dR, dS, aR, aS, otr = oracle.replay_launch(cfg, hand_table, golden_lut, R0, sig0, L, 1, L * n, n)
            for f in ("path_hash", "decisions", "updates", "terminal"):
                assert np.array_equal(tr[f], otr[f]), (f, L)
            assert np.array_equal(tr["utility"].view(np.uint32), otr["utility"].view(np.uint32)), L
            R1, S1, sig1 = ctx.download()
            eR, eS = helpers.expected_post(R0, S0, dR, dS, bool(cfg.use_cfr_plus))
            assert helpers.within_tol(R1, eR, aR).all() and helpers.within_tol(S1, eS, aS).all(), L
            assert np.array_equal(sig1, oracle.regret_matching(R1, row)), L


def test_train_graph_equals_launch_sequence(hand_table, golden_lut):
    """mccfr_train (captured hipGraph, device-advanced launch state) gives the
    tables of the same launches issued one by one and checked against the oracle."""
    cfg = mccfr.cfg_3p_20bb(table_log2=14, seed=11)
    n, iters = 2048, 3
    R_seq, S_seq, _ = run_stepwise(cfg, hand_table, golden_lut, n=n, launches=iters * cfg.num_players)
    with mccfr.Context(cfg, n) as ctx:
        ctx.set_hand_table(hand_table)
        ctx.train(iters, n)
        R, S, sig = ctx.download()
        assert ctx.counters() == (iters * cfg.num_players * n, iters * cfg.num_players, iters + 1)
    assert np.allclose(R, R_seq, rtol=1e-6, atol=1e-6)
    assert np.allclose(S, S_seq, rtol=1e-6, atol=1e-6)


def test_randomized_tables_follow_the_stream(hand_table):
    """mccfr_randomize_tables: regret[slot][a] = u01(draw a of stream (seed, slot, 'rand'))."""
    cfg = mccfr.cfg_6p_100bb(table_log2=12, seed=0)
    with mccfr.Context(cfg, 64) as ctx:
        ctx.randomize_tables(1234)
        R, S, sig = ctx.download()
    R = R.reshape(-1, 8)
    O = oracle.lib()
    for slot in (0, 1, 77, 4095):
        want = [O.orc_u01(O.orc_draw_u32(1234, slot, 0x72616E64, a)) for a in range(6)]
        assert np.array_equal(R[slot, :6], np.array(want, np.float32))
        assert (R[slot, 6:] == 0).all()
    assert (S == 0).all()
    assert np.array_equal(sig, oracle.regret_matching(R.reshape(-1), 8))


def test_full_size_6max_100bb_lanes_match_oracle(hand_table, golden_lut):
    """6 players, 100bb, 2^26-slot tables, 2^20 traversals per launch, from a
    randomized regret state: after warm-up launches, every lane of a launch is
    replayed by the oracle from the downloaded pre-launch tables: bit-exact;
    every touched table entry within tolerance of the replay's FP64 sums,
    every other entry unchanged; strategy-sum mass equals t * (update nodes)."""
    cfg = mccfr.cfg_6p_100bb(table_log2=26, seed=0)
    n = 1 << 20
    with mccfr.Context(cfg, n) as ctx:
        ctx.set_hand_table(hand_table)
        ctx.randomize_tables(5)
        tid = 0
        for L in range(6):  # one iteration, all six seats
            ctx.run_launch(L, 1, tid, n)
            tid += n
        R0, S0, sig0 = ctx.download()
        tr = ctx.run_launch(0, 2, tid, n, trace=True)
        recs, _, _ = ctx.last_launch_stats()
        R1, S1, _ = ctx.download()
        idx, dR, dS, aR, aS, otr = oracle.replay_sparse(cfg, hand_table, golden_lut, R0, sig0, 0, 2, tid, n)
        for f in ("path_hash", "decisions", "updates", "terminal"):
            assert np.array_equal(tr[f], otr[f]), f
        assert np.array_equal(tr["utility"].view(np.uint32), otr["utility"].view(np.uint32))
        mass = float(S1.astype(np.float64).sum() - S0.astype(np.float64).sum())
        nodes = int(tr["updates"].astype(np.int64).sum())
        assert recs == nodes  # every update node emitted once
        assert abs(mass - 2.0 * nodes) <= 1e-4 * 2.0 * nodes
        assert (R1 >= 0).all()
        helpers.check_launch_sparse(R0, S0, R1, S1, idx, dR, dS, aR, aS)
