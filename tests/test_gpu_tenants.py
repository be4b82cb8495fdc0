"""Tenants, batched launches, the multi-tenant training graph, the autotuner
and the re-solve server, on the device:

* postflop-root tenants against the oracle's FP64 replay, launch by launch
  (traces bit-exact; tables within the REL_TOL rule of helpers.py);
* a tenant's tables after launches shared with other tenants equal its tables
  after the same launches alone, bit for bit (sorted, fixed-order reductions);
* the captured multi-tenant graph equals the same launches issued one by one;
* results do not depend on the walk's launch shape;
* a re-solve gives the same tables alone, in a mix, and one-at-a-time.
"""
import numpy as np
import pytest

import helpers
import mccfr
import oracle
import workloads
from test_oracle_roots import fixed_roots

pytestmark = pytest.mark.gpu


def small_ref(N=3, stack=400, log2=12, seed=1):
    return mccfr.reference_cfg(num_players=N, starting_stack=stack, table_log2=log2, seed=seed)


def seats_of(cfg, root):
    if root is None:
        return list(range(cfg.num_players))
    return [p for p in range(cfg.num_players) if (int(root.active_mask) >> p) & 1]


@pytest.mark.parametrize("start", ["zero", "random"])
@pytest.mark.parametrize("k", range(4))
def test_postflop_root_stepwise_parity(k, start, hand_table, golden_lut):
    cfg, root = fixed_roots()[k]
    row = 8
    seats = seats_of(cfg, root)
    n = 3000
    with mccfr.Context(small_ref(), 4096) as ctx:
        ctx.set_hand_table(hand_table)
        t = ctx.tenant_create(cfg, root)
        if start == "random":
            ctx.tenant_randomize(t, 99)
        tid = 0
        for L in range(2 * len(seats)):
            up, w = seats[L % len(seats)], L // len(seats) + 1
            R0, S0, sig0 = ctx.tenant_download(t)
            assert np.array_equal(sig0, oracle.regret_matching(R0, row))
            print(f"root {k} launch {L}: up={up} t={w} tid={tid}", flush=True)
            tr = ctx.run_segments([(t, n, up, w, tid)], trace=True)
            dR, dS, aR, aS, otr = oracle.replay_launch(cfg, hand_table, golden_lut, R0, sig0, up, w, tid, n, root=root)
            for f in ("path_hash", "decisions", "updates", "terminal"):
                assert np.array_equal(tr[f], otr[f]), (f, L)
            assert np.array_equal(tr["utility"].view(np.uint32), otr["utility"].view(np.uint32)), L
            R1, S1, sig1 = ctx.tenant_download(t)
            eR, eS = helpers.expected_post(R0, S0, dR, dS, bool(cfg.use_cfr_plus))
            assert helpers.within_tol(R1, eR, aR).all(), L
            assert helpers.within_tol(S1, eS, aS).all(), L
            tid += n


def mix_games():
    """Preflop games of several sizes and menus plus postflop roots, all row 8."""
    games = [
        (mccfr.reference_cfg(num_players=2, starting_stack=1000, table_log2=13, seed=41), None, 1000),
        (mccfr.cfg_6p_100bb(table_log2=15, seed=42), None, 5000),
        (mccfr.cfg_3p_20bb(table_log2=12, seed=43), None, 300),
    ]
    for cfg, root in fixed_roots()[:3]:
        games.append((cfg, root, 777))
    return games


def run_mixed(hand_table, games, launches, together):
    """Runs `launches` rounds; in round L every game g runs its count for seat
    L mod seats with t = L // seats + 1.  Returns per-game (R, S) and traces."""
    with mccfr.Context(small_ref(), 16384) as ctx:
        ctx.set_hand_table(hand_table)
        ids = [ctx.tenant_create(cfg, root) for cfg, root, _ in games]
        for t in ids:
            ctx.tenant_randomize(t, 7 + t)
        tids = [0] * len(games)
        traces = [[] for _ in games]
        for L in range(launches):
            segs = []
            for g, ((cfg, root, n), t) in enumerate(zip(games, ids)):
                seats = seats_of(cfg, root)
                segs.append((t, n, seats[L % len(seats)], L // len(seats) + 1, tids[g]))
                tids[g] += n
            if together:
                tr = ctx.run_segments(segs, trace=True)
                off = 0
                for g, s in enumerate(segs):
                    traces[g].append(tr[off:off + s[1]])
                    off += s[1]
            else:
                for g, s in enumerate(segs):
                    traces[g].append(ctx.run_segments([s], trace=True))
        out = [ctx.tenant_download(t)[:2] for t in ids]
    return out, traces


def test_batched_launch_equals_each_tenant_alone(hand_table):
    games = mix_games()
    mixed, tr_m = run_mixed(hand_table, games, 4, together=True)
    alone, tr_a = run_mixed(hand_table, games, 4, together=False)
    for g in range(len(games)):
        for a, b in zip(tr_m[g], tr_a[g]):
            assert np.array_equal(a, b), g
        assert np.array_equal(mixed[g][0], alone[g][0]), g
        assert np.array_equal(mixed[g][1], alone[g][1]), g


def test_batched_launch_matches_oracle(hand_table, golden_lut):
    """One mixed launch checked per tenant against the oracle replay."""
    games = mix_games()
    with mccfr.Context(small_ref(), 16384) as ctx:
        ctx.set_hand_table(hand_table)
        ids = [ctx.tenant_create(cfg, root) for cfg, root, _ in games]
        for t in ids:
            ctx.tenant_randomize(t, 100 + t)
        pre = [ctx.tenant_download(t) for t in ids]
        segs = [(t, n, seats_of(cfg, root)[-1], 3, 1000 * t) for t, (cfg, root, n) in zip(ids, games)]
        tr = ctx.run_segments(segs, trace=True)
        off = 0
        for (t, n, up, w, tid), (cfg, root, _), (R0, S0, sig0) in zip(segs, games, pre):
            dR, dS, aR, aS, otr = oracle.replay_launch(cfg, hand_table, golden_lut, R0, sig0, up, w, tid, n,
                                                       root=root)
            mine = tr[off:off + n]
            off += n
            assert np.array_equal(mine["path_hash"], otr["path_hash"]), t
            assert np.array_equal(mine["utility"].view(np.uint32), otr["utility"].view(np.uint32)), t
            R1, S1, _ = ctx.tenant_download(t)
            eR, eS = helpers.expected_post(R0, S0, dR, dS, bool(cfg.use_cfr_plus))
            assert helpers.within_tol(R1, eR, aR).all(), t
            assert helpers.within_tol(S1, eS, aS).all(), t


def test_train_tenants_graph_equals_launch_sequence(hand_table):
    games = mix_games()[:4]
    counts = [n for _, _, n in games]
    launches = 7
    with mccfr.Context(small_ref(), 16384) as ctx:
        ctx.set_hand_table(hand_table)
        ids = [ctx.tenant_create(cfg, root) for cfg, root, _ in games]
        ctx.train_tenants(ids, counts, 3)
        ctx.train_tenants(ids, counts, launches - 3)  # continues every tenant's schedule
        graph = [ctx.tenant_download(t)[:2] for t in ids]
        ctrs = [ctx.tenant_counters(t) for t in ids]
    with mccfr.Context(small_ref(), 16384) as ctx:
        ctx.set_hand_table(hand_table)
        ids = [ctx.tenant_create(cfg, root) for cfg, root, _ in games]
        tids = [0] * len(games)
        for L in range(launches):
            segs = []
            for g, ((cfg, root, n), t) in enumerate(zip(games, ids)):
                seats = seats_of(cfg, root)
                it = L // len(seats) + 1
                segs.append((t, n, seats[L % len(seats)], it if cfg.use_linear_cfr else 1, tids[g]))
                tids[g] += n
            ctx.run_segments(segs)
        seq = [ctx.tenant_download(t)[:2] for t in ids]
    for g, (cfg, root, n) in enumerate(games):
        assert np.array_equal(graph[g][0], seq[g][0]), g
        assert np.array_equal(graph[g][1], seq[g][1]), g
        seats = seats_of(cfg, root)
        assert ctrs[g] == (launches * n, launches, launches // len(seats) + 1, seats[launches % len(seats)])


def test_graph_replays_one_at_a_time_with_eager_work_between(hand_table):
    """The captured training graph replayed launch by launch, with eager
    launches of another tenant, downloads and synchronizations between the
    replays: every replay's tables equal the eager sequence's after as many
    launches (bit for bit)."""
    games = mix_games()[:4]
    counts = [n for _, _, n in games]
    launches = 6
    other = mccfr.cfg_3p_20bb(table_log2=12, seed=99)
    graph_states, seq_states = [], []
    with mccfr.Context(small_ref(), 16384) as ctx:
        ctx.set_hand_table(hand_table)
        ids = [ctx.tenant_create(cfg, root) for cfg, root, _ in games]
        x = ctx.tenant_create(other)
        x_tid = 0
        for L in range(launches):
            ctx.train_tenants(ids, counts, 1)
            graph_states.append([ctx.tenant_download(t)[:2] for t in ids])
            ctx.run_segments([(x, 2048, L % 3, L // 3 + 1, x_tid)])  # eager work between replays
            x_tid += 2048
            ctx.tenant_normalized_strategy(x)
            ctx.sync()
    with mccfr.Context(small_ref(), 16384) as ctx:
        ctx.set_hand_table(hand_table)
        ids = [ctx.tenant_create(cfg, root) for cfg, root, _ in games]
        tids = [0] * len(games)
        for L in range(launches):
            segs = []
            for g, ((cfg, root, n), t) in enumerate(zip(games, ids)):
                seats = seats_of(cfg, root)
                it = L // len(seats) + 1
                segs.append((t, n, seats[L % len(seats)], it if cfg.use_linear_cfr else 1, tids[g]))
                tids[g] += n
            ctx.run_segments(segs)
            seq_states.append([ctx.tenant_download(t)[:2] for t in ids])
    for L in range(launches):
        for g in range(len(games)):
            assert np.array_equal(graph_states[L][g][0], seq_states[L][g][0]), (L, g)
            assert np.array_equal(graph_states[L][g][1], seq_states[L][g][1]), (L, g)


def test_results_do_not_depend_on_launch_shape(hand_table):
    cfg = mccfr.cfg_6p_100bb(table_log2=14, seed=3)
    out = []
    for wg, tpl in ((256, 1), (64, 2), (1024, 4), (128, 4)):
        with mccfr.Context(cfg, 8192) as ctx:
            ctx.set_hand_table(hand_table)
            ctx.set_launch_shape(wg, tpl)
            assert ctx.launch_shape() == (wg, tpl)
            ctx.randomize_tables(5)
            ctx.train(2, 8000)
            out.append(ctx.download()[:2])
    for R, S in out[1:]:
        assert np.array_equal(R, out[0][0])
        assert np.array_equal(S, out[0][1])


def test_persistent_pool_equals_one_traversal_per_lane(hand_table):
    """The persistent shape at 2^20 traversals: one workgroup per resident slot,
    so each wave's pool holds 2-3 traversals per lane (the last round is only
    partly filled) and finished lanes take the pool's next traversal.  Every
    lane's record and both tables equal the default shape's."""
    cfg = mccfr.cfg_6p_100bb(table_log2=16, seed=4)
    n = 1 << 20
    got = []
    for wg, tpl in ((256, 1), (256, 0), (64, 0)):
        with mccfr.Context(cfg, n) as ctx:
            ctx.set_hand_table(hand_table)
            ctx.set_launch_shape(wg, tpl)
            ctx.randomize_tables(6)
            ctx.run_launch(0, 1, 0, n)
            tr = ctx.run_launch(1, 1, n, n, trace=True)
            got.append((tr, ctx.download()[:2]))
    tr0, (R0, S0) = got[0]
    for tr, (R, S) in got[1:]:
        assert np.array_equal(tr, tr0)
        assert np.array_equal(R, R0) and np.array_equal(S, S0)


def test_autotune_picks_a_shape(hand_table):
    games, counts = workloads.uneven_mix(seed=2, n_games=8, table_log2=14, traversals=1 << 15)
    with mccfr.Context(small_ref(), 1 << 16) as ctx:
        ctx.set_hand_table(hand_table)
        ids = [ctx.tenant_create(g.cfg) for g in games]
        segs = [(t, int(c), 0, 1, 0) for t, c in zip(ids, counts)]
        before = [ctx.tenant_download(t)[0].copy() for t in ids]
        wg, tpl, ms = ctx.autotune(segs, reps=2)
        assert (wg, tpl) == ctx.launch_shape()
        assert wg in (64, 128, 256, 512, 1024) and tpl in (1, 2, 4) and ms > 0
        for t, R in zip(ids, before):  # the autotuner only walks: tables untouched
            assert np.array_equal(ctx.tenant_download(t)[0], R)


def test_upload_with_negative_regrets_is_clamped_after_the_next_launch(hand_table, golden_lut):
    """CFR+ clamps every slot after every launch (cfr_gpu.cu:1018-1022): after an
    upload holding negative regrets the next launch leaves none, touched or not."""
    cfg = small_ref(N=2, stack=1000, log2=12, seed=9)
    slots, row = 1 << 12, 8
    rng = np.random.default_rng(0)
    R = (rng.standard_normal(slots * row) * 10).astype(np.float32)
    with mccfr.Context(cfg, 2048) as ctx:
        ctx.set_hand_table(hand_table)
        ctx.upload(R, np.zeros_like(R))
        R0, S0, sig0 = ctx.download()
        ctx.run_launch(0, 1, 0, 2000)
        dR, dS, aR, aS, _ = oracle.replay_launch(cfg, hand_table, golden_lut, R0, sig0, 0, 1, 0, 2000)
        R1, S1, _ = ctx.download()
    assert (R1 >= 0).all()
    eR, eS = helpers.expected_post(R0, S0, dR, dS, True)
    assert helpers.within_tol(R1, eR, aR).all()


def serve(hand_table, reqs, max_running, budget=1 << 31):
    """Runs requests through a server (all submitted up front); returns
    {index: (R, S)} and the order requests finished in."""
    out, order = {}, []
    with mccfr.Context(small_ref(), 1 << 15) as ctx:
        ctx.set_hand_table(hand_table)
        with mccfr.Server(ctx, budget, max_running) as srv:
            ids = [srv.submit(r.cfg, r.root, r.iterations, r.batch) for r in reqs]
            pending = set(range(len(reqs)))
            max_seen = 0
            for _ in range(100000):
                running, waiting = srv.step(wait=True)
                max_seen = max(max_seen, running)
                for i in sorted(pending):
                    st, _ = srv.poll(ids[i])
                    if st == mccfr.DONE:
                        t = srv.tenant(ids[i])
                        out[i] = ctx.tenant_download(t)[:2]
                        strat = srv.result(ids[i])
                        assert np.isfinite(strat).all()
                        srv.release(ids[i])
                        pending.discard(i)
                        order.append(i)
                if not pending:
                    break
            assert not pending
            assert max_seen <= max_running
    serve.max_seen = max_seen
    return out, order


def small_requests(n):
    reqs = workloads.resolve_requests(n, min_iters=200, max_iters=3000)
    for i, r in enumerate(reqs):
        r.cfg.table_log2 = 14 + (i % 3)
        r.batch = workloads.request_batch(r.iterations)
    return reqs


def test_server_result_is_the_same_alone_in_the_mix_and_one_at_a_time(hand_table):
    reqs = small_requests(6)
    mix, _ = serve(hand_table, reqs, max_running=6)
    one, order = serve(hand_table, reqs, max_running=1)
    assert order == list(range(6))  # first come, first served
    for i in range(6):
        alone, _ = serve(hand_table, [reqs[i]], max_running=1)
        for k in range(2):
            assert np.array_equal(mix[i][k], alone[0][k]), (i, k)
            assert np.array_equal(one[i][k], alone[0][k]), (i, k)
        assert np.count_nonzero(mix[i][1]) > 0


def test_server_admits_by_hbm_budget(hand_table):
    reqs = small_requests(5)
    for r in reqs:
        r.cfg.table_log2 = 16
    fp = mccfr.tenant_footprint(reqs[0].cfg)
    out, order = serve(hand_table, reqs, max_running=8, budget=2 * fp + 4096)
    assert sorted(order) == list(range(5))
    assert serve.max_seen == 2  # two tables fit the pool at a time


def test_server_admits_requests_that_arrive_while_others_run(hand_table):
    """Requests submitted between steps while others are on the device, with
    two running places and a pool for two tables, so later ones wait for a
    place and for memory: each finishes with the tables it gets alone."""
    reqs = small_requests(7)
    for r in reqs:
        r.cfg.table_log2 = 15
    fp = mccfr.tenant_footprint(reqs[0].cfg)
    got = {}
    submitted_mid_flight = waited = 0
    with mccfr.Context(small_ref(), 1 << 15) as ctx:
        ctx.set_hand_table(hand_table)
        with mccfr.Server(ctx, 2 * fp + 4096, 2) as srv:
            ids = {}
            pending = set()
            nxt = 0
            for step in range(100000):
                if nxt < len(reqs) and step % 3 == 0:  # one arrival every third step
                    r = reqs[nxt]
                    running_before = sum(srv.poll(ids[i])[0] == mccfr.RUNNING for i in pending)
                    submitted_mid_flight += running_before > 0
                    ids[nxt] = srv.submit(r.cfg, r.root, r.iterations, r.batch)
                    pending.add(nxt)
                    nxt += 1
                running, waiting = srv.step(wait=True)
                assert running <= 2
                waited += waiting > 0
                for i in sorted(pending):
                    if srv.poll(ids[i])[0] == mccfr.DONE:
                        got[i] = ctx.tenant_download(srv.tenant(ids[i]))[:2]
                        srv.release(ids[i])
                        pending.discard(i)
                if nxt == len(reqs) and not pending:
                    break
    assert submitted_mid_flight >= 3 and waited > 0
    for i, r in enumerate(reqs):
        alone, _ = serve(hand_table, [r], max_running=1)
        for k in range(2):
            assert np.array_equal(got[i][k].view(np.uint32), alone[0][k].view(np.uint32)), (i, k)


def test_server_fails_a_request_whose_launch_outgrows_the_scratch(hand_table):
    """After mccfr_set_max_batch shrinks the context below a running request's
    launch, the step reports MCCFR_E_STATE, the request ends as failed (state
    4) and can be released, and the server keeps serving others."""
    reqs = small_requests(2)
    with mccfr.Context(small_ref(), 4096) as ctx:
        ctx.set_hand_table(hand_table)
        with mccfr.Server(ctx, 1 << 30, 4) as srv:
            big = srv.submit(reqs[0].cfg, reqs[0].root, 100000, 4096)
            srv.step(wait=False)
            assert srv.poll(big)[0] == mccfr.RUNNING
            ctx.set_max_batch(1024)
            with pytest.raises(mccfr.MccfrError) as e:
                for _ in range(4):
                    srv.step(wait=True)
            assert e.value.code == mccfr.E_STATE
            assert srv.poll(big)[0] == 4
            srv.release(big)
            ok = srv.submit(reqs[1].cfg, reqs[1].root, 2000, 512)
            for _ in range(10000):
                srv.step(wait=True)
                if srv.poll(ok)[0] == mccfr.DONE:
                    break
            assert srv.poll(ok)[0] == mccfr.DONE
