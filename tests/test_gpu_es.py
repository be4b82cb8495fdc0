"""External sampling (MCCFR_SAMPLE_EXTERNAL) on the device against the oracle:
every traversal's record bit-exact (root value, path hash over every entered
(node, action), node, leaf and legal-action counts), tables within the
tolerance of the FP64 replay after every launch, from zero and from random
tables, preflop and postflop roots; the training graph against the launch
sequence; external-sampling tenants batched together against each alone; and
the refusals (a launch whose records exceed the scratch, mixing samplings)."""
import numpy as np
import pytest

import helpers
import mccfr
import oracle

pytestmark = pytest.mark.gpu
FIELDS = ("path_hash", "decisions", "updates", "terminal", "update_legal", "reserved")


def es(cfg):
    cfg.sampling = mccfr.SAMPLE_EXTERNAL
    return cfg


CONFIGS = {
    "3p_20bb": lambda: es(mccfr.cfg_3p_20bb(table_log2=14, seed=3)),
    "6max_100bb": lambda: es(mccfr.cfg_6p_100bb(table_log2=15, seed=4)),
    "2p_ref_10bb": lambda: es(mccfr.reference_cfg(num_players=2, starting_stack=200, table_log2=14, seed=5)),
    "6p_200bb_10": lambda: es(mccfr.cfg_6p_200bb_10(table_log2=14, seed=6)),
}


def check_launch(ctx, cfg, hr, lut, up, t, tid, n, root=None, tenant=0):
    R0, S0, sig0 = ctx.tenant_download(tenant)
    if tenant == 0 and root is None:
        tr = ctx.run_launch(up, t, tid, n, trace=True)
    else:
        tr = ctx.run_segments([(tenant, n, up, t, tid)], trace=True)
    dR, dS, aR, aS, otr = oracle.replay_launch(cfg, hr, lut, R0, sig0, up, t, tid, n, root=root)
    for f in FIELDS:
        assert np.array_equal(tr[f], otr[f]), f
    assert np.array_equal(tr["utility"].view(np.uint32), otr["utility"].view(np.uint32))
    R1, S1, sig1 = ctx.tenant_download(tenant)
    eR, eS = helpers.expected_post(R0, S0, dR, dS, bool(cfg.use_cfr_plus))
    assert helpers.within_tol(R1, eR, aR).all()
    assert helpers.within_tol(S1, eS, aS).all()
    row = 8 if cfg.num_actions <= 8 else 16
    assert np.array_equal(sig1, oracle.regret_matching(R1, row))
    return tr


@pytest.mark.parametrize("name", sorted(CONFIGS))
@pytest.mark.parametrize("start", ["zero", "random"])
def test_es_launches_match_oracle(hand_table, golden_lut, name, start):
    cfg = CONFIGS[name]()
    n = 1024
    with mccfr.Context(cfg, 64 * n) as ctx:  # records: 12-540 per traversal here (20 lanes of scratch each)
        ctx.set_hand_table(hand_table)
        if start == "random":
            ctx.randomize_tables(9)
        tid = 0
        for L in range(2 * cfg.num_players):
            up, t = L % cfg.num_players, L // cfg.num_players + 1
            print(f"es {name} {start}: launch {L}", flush=True)
            tr = check_launch(ctx, cfg, hand_table, golden_lut, up, t, tid, n)
            tid += n
        assert tr["updates"].astype(np.int64).sum() > n  # whole trees, not paths


def test_es_postflop_roots_match_oracle(hand_table, golden_lut):
    cfg = es(mccfr.cfg_6p_100bb(table_log2=14, seed=8))
    board = [12, 25, 38, 3, 17]  # A A A 5 5 across suits
    roots = [mccfr.root_state(1, 2, [0, 1, 3, 5], [1500, 900, 2000, 1800, 2000, 400], [100, 100, 0, 100, 0, 100],
                              board[:3]),
             mccfr.root_state(3, 0, [1, 2], [0, 700, 1200, 0, 0, 0], [20, 300, 300, 10, 0, 0], board)]
    with mccfr.Context(mccfr.cfg_6p_100bb(table_log2=4), 8192) as ctx:
        ctx.set_hand_table(hand_table)
        for k, root in enumerate(roots):
            t = ctx.tenant_create(cfg, root)
            ctx.tenant_randomize(t, 3 + k)
            seats = [p for p in range(6) if root.active_mask >> p & 1]
            for L, up in enumerate(seats):
                check_launch(ctx, cfg, hand_table, golden_lut, up, 1, 512 * L, 512, root=root, tenant=t)


def test_es_train_graph_equals_launch_sequence(hand_table, golden_lut):
    cfg = es(mccfr.cfg_3p_20bb(table_log2=13, seed=12))
    n, iters = 1024, 2
    with mccfr.Context(cfg, 16 * n) as a, mccfr.Context(cfg, 16 * n) as b:
        for c in (a, b):
            c.set_hand_table(hand_table)
        tid = 0
        for L in range(iters * cfg.num_players):
            a.run_launch(L % cfg.num_players, L // cfg.num_players + 1, tid, n)
            tid += n
        b.train(iters, n)
        Ra, Sa, _ = a.download()
        Rb, Sb, _ = b.download()
    assert np.allclose(Ra, Rb, rtol=1e-6, atol=1e-5)
    assert np.allclose(Sa, Sb, rtol=1e-6, atol=1e-5)
    assert np.count_nonzero(Sa) > 100


def test_es_tenants_batched_equal_each_alone(hand_table, golden_lut):
    cfgs = [es(mccfr.cfg_3p_20bb(table_log2=12, seed=20 + i)) for i in range(3)]
    cfgs.append(es(mccfr.cfg_6p_100bb(table_log2=12, seed=30)))
    n = 512
    with mccfr.Context(mccfr.cfg_3p_20bb(table_log2=4), 32 * n) as ctx:
        ctx.set_hand_table(hand_table)
        ids = [ctx.tenant_create(c) for c in cfgs]
        for t in ids:
            ctx.tenant_randomize(t, t)
        pre = [ctx.tenant_download(t) for t in ids]
        ctx.run_segments([(t, n, 1, 2, 100) for t in ids])
        mixed = [ctx.tenant_download(t)[:2] for t in ids]
        for t, c, (R0, S0, sig0), (Rm, Sm) in zip(ids, cfgs, pre, mixed):
            ctx.tenant_upload(t, R0, S0)
            ctx.run_segments([(t, n, 1, 2, 100)])
            Ra, Sa, _ = ctx.tenant_download(t)
            assert np.allclose(Rm, Ra, rtol=1e-6, atol=1e-5)
            assert np.allclose(Sm, Sa, rtol=1e-6, atol=1e-5)
            dR, dS, aR, aS, _ = oracle.replay_launch(c, hand_table, golden_lut, R0, sig0, 1, 2, 100, n)
            eR, eS = helpers.expected_post(R0, S0, dR, dS)
            assert helpers.within_tol(Rm, eR, aR).all() and helpers.within_tol(Sm, eS, aS).all()
        # a launch may not mix the two samplings
        os_t = ctx.tenant_create(mccfr.cfg_3p_20bb(table_log2=12, seed=1))
        with pytest.raises(mccfr.MccfrError) as e:
            ctx.run_segments([(ids[0], n, 0, 1, 0), (os_t, n, 0, 1, 0)])
        assert e.value.code == mccfr.E_ARG


def test_es_launch_beyond_the_record_scratch_is_refused(hand_table):
    """Heads-up with deep stacks and the 8-action menu: whole trees hold
    hundreds to thousands of records, far beyond 20 per lane of a 256-lane
    scratch; the launch reports MCCFR_E_STATE and leaves the tables as they
    were, and the context keeps working."""
    cfg = es(mccfr.reference_cfg(num_players=2, starting_stack=1000, table_log2=12, seed=1))
    with mccfr.Context(cfg, 256) as ctx:
        ctx.set_hand_table(hand_table)
        ctx.randomize_tables(4)
        R0, S0, _ = ctx.download()
        with pytest.raises(mccfr.MccfrError) as e:
            ctx.run_launch(0, 1, 0, 256)
        assert e.value.code == mccfr.E_STATE
        R1, S1, _ = ctx.download()
        assert np.array_equal(R0, R1) and np.array_equal(S0, S1)
        with pytest.raises(mccfr.MccfrError):
            ctx.run_launch_noalloc(0, 1, 0, 1)  # external sampling needs the context's scratch
        ctx.set_max_batch(1 << 18)
        ctx.run_launch(0, 1, 0, 1)  # one tree fits now
        R2, _, _ = ctx.download()
        assert not np.array_equal(R0, R2)


def test_es_depth_first_and_generation_walks_agree(hand_table, golden_lut):
    """A launch of several external-sampling segments takes the depth-first
    walk (k_walk_es), a launch of one segment the generation walk (k_es_gen /
    k_es_back).  Both explore the same tasks with the same streams: every
    traversal's record is bit-exact between the two and against the oracle."""
    cfgs = [es(mccfr.cfg_3p_20bb(table_log2=12, seed=40)), es(mccfr.cfg_6p_100bb(table_log2=12, seed=41)),
            es(mccfr.reference_cfg(num_players=2, starting_stack=200, table_log2=12, seed=42))]
    n = 512
    with mccfr.Context(mccfr.cfg_3p_20bb(table_log2=4), 64 * n) as ctx:
        ctx.set_hand_table(hand_table)
        ids = [ctx.tenant_create(c) for c in cfgs]
        for t in ids:
            ctx.tenant_randomize(t, 50 + t)
        pre = [ctx.tenant_download(t) for t in ids]
        segs = [(t, n, 1, 2, 1000 * (k + 1)) for k, t in enumerate(ids)]
        mixed = ctx.run_segments(segs, trace=True)  # depth first
        for k, (t, c, (R0, S0, sig0)) in enumerate(zip(ids, cfgs, pre)):
            part = mixed[k * n:(k + 1) * n]
            _, _, _, _, otr = oracle.replay_launch(c, hand_table, golden_lut, R0, sig0, 1, 2, 1000 * (k + 1), n)
            for f in FIELDS:
                assert np.array_equal(part[f], otr[f]), (k, f)
            assert np.array_equal(part["utility"].view(np.uint32), otr["utility"].view(np.uint32))
            ctx.tenant_upload(t, R0, S0)
            alone = ctx.run_segments([segs[k]], trace=True)  # by generations
            for f in FIELDS + ("utility",):
                assert np.array_equal(alone[f].view(np.uint32) if f == "utility" else alone[f],
                                      part[f].view(np.uint32) if f == "utility" else part[f]), (k, f)
            assert alone["updates"].astype(np.int64).sum() > n
