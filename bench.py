#!/usr/bin/env python3
"""Benchmark of the MCCFR hot path on one MI355X.

Headline workload (`value`): 6-max 100bb NLH -- 6 players, 2,000-chip stacks
at blinds 10/20, menu {fold, check/call, 1/2 pot, pot, 2 pot, all-in}, at most
3 raises per street; FP32 regret and strategy-sum tables of 2^26 slots; one
step = one launch of 2^20 traversals (deal, buckets, betting-tree walk, record
sort, owner-computes CFR+ table update), the update player cycling over the
seats as in the training loop.  Tables and the hand table are resident in HBM
before timing.

The same JSON line carries:
  roofline      the walk kernel (the launch's dominant kernel): algorithmic
                bytes per launch / its mean duration (HIP events around it on
                the library stream), HBM bytes per launch from the committed
                rocprofv3 FETCH_SIZE/WRITE_SIZE summary;
  batched       32 uneven games (2-6 players, 10-200bb, seed 2) sharing
                2^21-traversal launches from one captured graph (autotuned
                shape) vs the same games each launched alone;
  serving       re-solve requests arriving Poisson at 200/s (seeds 3 and 4):
                p50/p99 latency from arrival to final tables, aggregate
                traversals/s; and all requests at once, served vs one at a time;
  hand_eval     2^28 uniform 7-card hands through both evaluators;
  cpu_baseline  the oracle's FP32 restatement of the reference launch on a
                bounded sample, on the process's share of the host cores and
                on one thread, plus the reference's table walk over the hand
                microbenchmark's first 2^24 hands on the host;
  config1       the reference's CPU-runnable case (3 players, 20bb, 2^22
                slots, launches of 2^10) on the CPU and through the GPU graph;
  large_table   12 launches of 2^22 traversals against 2^28-slot 16-wide
                tables (the large-table stress at size, fewer traversals);
  external_sampling  the headline game with every legal action explored at
                update-player nodes (launches of 2^20 traversals).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallel-mccfr-for-multiplayer-nlh_amd"))

METRIC = "MCCFR traversals/s (6-max 100bb) + p99 re-solve latency, 1 MI355X"
HBM_PEAK_GBS = 8000.0
HAND_TAG = 0x68616E64


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def walk_bytes(tr, num_players, row):
    """Algorithmic bytes of the walk kernel for one launch (per-traversal figure
    x traversals): root records in (4 B buckets + 2 B rank per seat, 1 B
    dealer), one BF16 strategy row per decision node (2 B x padded row), out:
    one 8-byte record (key, action word) per update-player node and the
    traversal's 1-byte node count and 4-byte utility."""
    n = len(tr)
    d = tr["decisions"].astype(np.float64).sum()
    u = tr["updates"].astype(np.float64).sum()
    return (6.0 * num_players + 1.0) * n + 2.0 * row * d + 8.0 * u + 5.0 * n


def path_bytes(tr, num_players, row):
    """The reference path's algorithmic bytes per launch (72 N hand-table
    words per traversal, a BF16 row per decision, an FP32 regret row and a
    regret/strategy-sum read-modify-write per legal action at update nodes)."""
    d = tr["decisions"].astype(np.float64).sum()
    u = tr["updates"].astype(np.float64).sum()
    v = tr["update_legal"].astype(np.float64).sum()
    return 72.0 * num_players * len(tr) + 2.0 * row * d + 4.0 * row * u + 16.0 * v


def hand_bench(ctx, log2n, reps=3):
    n = 1 << log2n
    dc = ctx.dalloc(n * 8)
    dr = ctx.dalloc(n * 2)
    out = {}
    try:
        ctx.random_hands_device(1, HAND_TAG, dc, n)
        ctx.sync()
        for ev, name in ((0, "table_walk"), (1, "arithmetic")):
            ctx.eval7_device(dc, 8, dr, n, ev)  # warm-up
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                ctx.eval7_device(dc, 8, dr, n, ev)
            ctx.sync()
            dt = (time.perf_counter() - t0) / reps
            out[name] = {"hands_per_s": n / dt, "ms": dt * 1e3, "gbs_37B_per_hand": 37.0 * n / dt / 1e9}
    finally:
        ctx.dfree(dc)
        ctx.dfree(dr)
    out["hands"] = n
    out["seed"] = 1
    return out


def host_cpus():
    """(threads the CPU legs use, CPUs this process may run on).  The GPU box
    gives a GPU's process a share of its host cores (OMP_NUM_THREADS there,
    16); the legs use that share, capped by the CPUs visible to the process."""
    visible = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(share, visible) if share > 0 else visible), visible


def cpu_baseline(mccfr, ctx, cfg, launch, tid, batch, launches, threads, launches_1=2):
    """The oracle's FP32 restatement of the reference launch on the GPU's own
    table state: on `threads` threads over `launches` launches, and on one
    thread over `launches_1` launches (bounded samples)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    O = oracle.lib()
    R0, S0, _ = ctx.download(layout=1)  # same table state as the GPU, action-major like the reference
    hr = mccfr.generate_hand_table()
    lut = mccfr.preflop_lut()
    N = cfg.num_players
    out = {}
    for key, th, nl in (("threads_all", threads, launches), ("threads_1", 1, launches_1)):
        R, S = R0.copy(), S0.copy()
        sig = np.zeros(R.size, np.uint16)
        secs = O.orc_cpu_baseline(C.byref(cfg), oracle.p(hr), oracle.p(lut), oracle.p(R), oracle.p(S),
                                  oracle.p(sig), nl, batch, launch % N, launch // N + 1, tid, th)
        if secs <= 0:
            return None
        out[key] = {"traversals_per_s": nl * batch / secs, "threads": th, "launches": nl, "seconds": secs}
    return out


def hand_eval_cpu(mccfr, ctx, log2n, threads):
    """The reference's table walk (hand_eval.cpp:134-143, the oracle's
    restatement) on the host over the first 2^log2n hands of the GPU
    microbenchmark's inputs (seed 1), on `threads` threads and on one."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    n = 1 << log2n
    dc = ctx.dalloc(n * 8)
    try:
        ctx.random_hands_device(1, HAND_TAG, dc, n)
        cards = np.empty(n * 8, np.uint8)
        ctx.d2h(cards, dc)
    finally:
        ctx.dfree(dc)
    hr = mccfr.generate_hand_table()
    out = {"hands": n, "seed": 1}
    for key, th in (("threads_all", threads), ("threads_1", 1)):
        ranks = np.empty(n, np.uint16)
        t0 = time.perf_counter()
        oracle.lib().orc_walk7_batch(oracle.p(hr), oracle.p(cards), 8, oracle.p(ranks), n, th)
        dt = time.perf_counter() - t0
        out[key] = {"hands_per_s": n / dt, "threads": th}
    return out


def config1_bench(mccfr, hr, threads, cpu_launches_all=100, cpu_launches_1=30, gpu_iters=200):
    """The reference's CPU-runnable case: 3 players, 20bb (400 chips), {F, C/C,
    pot, all-in}, <= 2 raises per street, 2^22-slot tables, launches of 2^10
    traversals from fresh (zero) tables and seed 0.  CPU: the oracle's
    restatement (full sweeps per launch, as the reference) on a bounded sample
    of the run's first launches; GPU: the captured training graph."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cfg = mccfr.cfg_3p_20bb(table_log2=22, seed=0)
    B = 1 << 10
    lut = mccfr.preflop_lut()
    n = (1 << 22) * 8
    out = {"workload": "config1: 3 players, 400-chip stacks, {F, C/C, pot, AI}, <=2 raises/street, 2^22 slots, "
                       "batch 2^10, fresh tables, seed 0"}
    for key, th, nl in (("cpu_threads_all", threads, cpu_launches_all), ("cpu_threads_1", 1, cpu_launches_1)):
        R = np.zeros(n, np.float32)
        S = np.zeros(n, np.float32)
        sig = np.zeros(n, np.uint16)
        secs = oracle.lib().orc_cpu_baseline(C.byref(cfg), oracle.p(hr), oracle.p(lut), oracle.p(R), oracle.p(S),
                                             oracle.p(sig), nl, B, 0, 1, 0, th)
        out[key] = {"traversals_per_s": nl * B / secs, "threads": th, "launches": nl}
    with mccfr.Context(cfg, B) as ctx:
        ctx.set_hand_table(hr)
        ctx.train(2, B)  # capture + warm-up
        ctx.sync()
        t0 = time.perf_counter()
        ctx.train(gpu_iters, B)
        ctx.sync()
        dt = time.perf_counter() - t0
    out["gpu"] = {"traversals_per_s": gpu_iters * cfg.num_players * B / dt, "launches": gpu_iters * cfg.num_players,
                  "ms_per_launch": dt * 1e3 / (gpu_iters * cfg.num_players)}
    return out


def harness_bench(mccfr, ctx, cfg, n_players=6, opponent=3, log2_hands=22, cpu_hands=100_000):
    """The strategy-quality harness (evaluate_strategy_np, eval.cpp:581-627) on the
    trained tables: device hands/s, and the oracle's hands/s on one thread with the
    reference's own stream (std::mt19937, one sequence) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    n = 1 << log2_hands
    ctx.evaluate_strategy(n_players, opponent, 1 << 16, seed=1)  # builds the harness buckets
    ctx.sync()
    t0 = time.perf_counter()
    res = ctx.evaluate_strategy(n_players, opponent, n, seed=42)
    gpu_s = time.perf_counter() - t0
    P = ctx.normalized_strategy().reshape(-1, ctx.row)
    hr = mccfr.generate_hand_table()
    lut = mccfr.preflop_lut()
    t0 = time.perf_counter()
    ref = oracle.eval_hands(cfg, hr, lut, P, n_players, opponent, cpu_hands, seed=42, stream=1)
    cpu_s = time.perf_counter() - t0
    return {"hands": n, "players": n_players, "opponent": "balanced", "gpu_hands_per_s": n / gpu_s,
            "gpu_ms": gpu_s * 1e3, "bb_per_100": res.bb_per_100,
            "cpu_reference_stream": {"hands_per_s": cpu_hands / cpu_s, "cores": 1, "hands": cpu_hands,
                                     "bb_per_100": float(ref.astype(np.float64).sum()) / cfg.big_blind
                                     / cpu_hands * 100.0}}


def batched_bench(mccfr, workloads, hr, launches, n_games=32, table_log2=20, total=1 << 21):
    """Uneven games in shared launches vs each game in its own context."""
    games, counts = workloads.uneven_mix(seed=2, n_games=n_games, table_log2=table_log2, traversals=total)
    lanes = int(sum((int(c) + 255) // 256 * 256 for c in counts))
    small = mccfr.reference_cfg(num_players=2, starting_stack=1000, table_log2=4)
    out = {"games": n_games, "traversals_per_launch": total, "table_slots_per_game": 1 << table_log2,
           "seed": 2, "launches": launches}
    with mccfr.Context(small, lanes) as ctx:
        ctx.set_hand_table(hr)
        ids = [ctx.tenant_create(g.cfg) for g in games]
        segs = [(t, int(c), 0, 1, 0) for t, c in zip(ids, counts)]
        wg, tpl, ms = ctx.autotune(segs, reps=2)
        out["autotuned_shape"] = {"walk_wg": wg, "traversals_per_lane": tpl, "walk_ms": ms}
        ctx.train_tenants(ids, counts, 2)  # capture + warm-up
        t0 = time.perf_counter()
        ctx.train_tenants(ids, counts, launches)
        dt = time.perf_counter() - t0
    out["batched_traversals_per_s"] = launches * total / dt
    ctxs = []
    try:
        for g, c in zip(games, counts):
            cx = mccfr.Context(g.cfg, int(c))
            cx.set_hand_table(hr)
            cx.train_tenants([0], [int(c)], 1)  # capture + warm-up
            ctxs.append(cx)
        t0 = time.perf_counter()
        for cx, c in zip(ctxs, counts):
            cx.train_tenants([0], [int(c)], launches)
        dt = time.perf_counter() - t0
    finally:
        for cx in ctxs:
            cx.close()
    out["plain_traversals_per_s"] = launches * total / dt
    out["speedup"] = out["batched_traversals_per_s"] / out["plain_traversals_per_s"]
    return out


def serve_run(mccfr, hr, reqs, budget, max_running, poisson):
    """Runs requests through a server; with poisson=True each is submitted at
    its arrival time, else all at once.  Returns (latencies s, traversals, wall s)."""
    small = mccfr.reference_cfg(num_players=2, starting_stack=1000, table_log2=4)
    lat, trav = [], 0
    with mccfr.Context(small, 1 << 20) as ctx:
        ctx.set_hand_table(hr)
        with mccfr.Server(ctx, budget, max_running) as srv:
            n, i = len(reqs), 0
            pending = {}
            t0 = time.perf_counter()
            while len(lat) < n:
                now = time.perf_counter() - t0
                while i < n and (not poisson or reqs[i].arrival <= now):
                    r = reqs[i]
                    rid = srv.submit(r.cfg, r.root, r.iterations, r.batch)
                    pending[rid] = r.arrival if poisson else 0.0
                    i += 1
                running, waiting = srv.step(wait=True)
                now = time.perf_counter() - t0
                for rid in list(pending):
                    st, tr = srv.poll(rid)
                    if st == mccfr.DONE:
                        lat.append(now - pending.pop(rid))
                        trav += tr
                        srv.release(rid)
                if running == 0 and waiting == 0 and i < n and poisson:
                    gap = reqs[i].arrival - (time.perf_counter() - t0)
                    if gap > 0:
                        time.sleep(min(gap, 0.002))
            wall = time.perf_counter() - t0
    return np.array(lat), trav, wall


def plain_graph_run(mccfr, hr, reqs):
    """Each request alone as plain per-request launches: its own tenant
    (tables allocated for it), its launches replayed from a captured training
    graph (batch traversals per launch for every active seat, rounds until the
    iteration budget is covered), then freed.  Returns (traversals, wall s)."""
    small = mccfr.reference_cfg(num_players=2, starting_stack=1000, table_log2=4)
    trav = 0
    with mccfr.Context(small, 1 << 20) as ctx:
        ctx.set_hand_table(hr)
        ctx.sync()
        t0 = time.perf_counter()
        for r in reqs:
            t = ctx.tenant_create(r.cfg, r.root)
            seats = bin(r.root.active_mask).count("1") if r.root is not None and r.root.street > 0 \
                else r.cfg.num_players
            launches = -(-r.iterations // r.batch) * seats
            ctx.train_tenants([t], [r.batch], launches)
            ctx.sync()
            ctx.tenant_destroy(t)
            trav += launches * r.batch
        wall = time.perf_counter() - t0
    return trav, wall


def serving_bench(mccfr, workloads, hr, n_req, sat_req, budget, max_running):
    reqs = workloads.resolve_requests(n_req, rate=200.0, arrival_seed=3, state_seed=4)
    lat, trav, wall = serve_run(mccfr, hr, reqs, budget, max_running, poisson=True)
    out = {"requests": n_req, "arrival_rate_per_s": 200.0, "seeds": {"arrivals": 3, "states": 4},
           "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
           "max_ms": float(lat.max() * 1e3), "traversals_per_s": trav / wall, "wall_s": wall,
           "iterations_total": int(sum(r.iterations for r in reqs)),
           "hbm_budget_bytes": budget, "max_running": max_running}
    sat = reqs[:sat_req]
    _, trav_b, wall_b = serve_run(mccfr, hr, sat, budget, max_running, poisson=False)
    _, trav_p, wall_p = serve_run(mccfr, hr, sat, budget, 1, poisson=False)
    trav_g, wall_g = plain_graph_run(mccfr, hr, sat)
    out["saturated"] = {"requests": sat_req, "served_traversals_per_s": trav_b / wall_b,
                        "one_at_a_time_traversals_per_s": trav_p / wall_p,
                        "speedup": (trav_b / wall_b) / (trav_p / wall_p),
                        "per_request_graph_traversals_per_s": trav_g / wall_g,
                        "speedup_vs_per_request_graph": (trav_b / wall_b) / (trav_g / wall_g)}
    return out


def es_bench(mccfr, batch_log2=20, table_log2=26, warmup=6, steps=12, table_seed=2026):
    """External sampling (every legal action at update-player nodes) on the
    headline game: 6 players, 100bb, the 6-action menu, <= 3 raises per
    street, 2^26-slot tables, launches of 2^batch_log2 traversals from a
    seeded random regret state.  The walk's algorithmic bytes sum, over each
    explored tree, the root record (6N + 1 B), a BF16 snapshot row per
    sampled (other seats') node, an FP32 regret row per update-player node
    and, per update-player node, its 12-byte sort record and its FP32 row of
    regret deltas."""
    cfg = mccfr.cfg_6p_100bb(table_log2=table_log2, seed=0)
    cfg.sampling = mccfr.SAMPLE_EXTERNAL
    B = 1 << batch_log2
    N = cfg.num_players
    # one record per update-player node, ~22 per traversal here (more as the
    # strategy sharpens); the sort input holds 20 per lane of max_batch
    with mccfr.Context(cfg, 4 * B) as ctx:
        ctx.use_generated_hand_table()
        ctx.randomize_tables(table_seed)
        row = ctx.row
        L = 0
        for _ in range(warmup):
            ctx.run_launch(L % N, L // N + 1, L * B, B)
            L += 1
        ctx.set_timing(True)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.run_launch(L % N, L // N + 1, L * B, B)
            L += 1
        ctx.sync()
        dt = time.perf_counter() - t0
        deal_ms, walk_ms, upd_ms = ctx.kernel_times(steps)
        ctx.set_timing(False)
        tr = ctx.run_launch(L % N, L // N + 1, L * B, B, trace=True)
        recs, long_runs, pieces = ctx.last_launch_stats()
    d = tr["decisions"].astype(np.float64).sum()
    u = tr["updates"].astype(np.float64).sum()
    v = tr["update_legal"].astype(np.float64).sum()
    walk_bytes = (6.0 * N + 1.0) * B + 2.0 * row * (d - u) + 4.0 * row * u + (12.0 + 4.0 * row) * u
    tree_bytes = 72.0 * N * B + 2.0 * row * d + u * 4.0 * row + 16.0 * v
    wms = float(walk_ms.mean())
    return {"workload": "6max_100bb, external sampling at update-player nodes", "traversals_per_launch": B,
            "table_slots": 1 << table_log2, "steps": steps, "traversals_per_s": steps * B / dt,
            "ms_per_launch": dt * 1e3 / steps,
            "kernel_ms": {"deal": float(deal_ms.mean()), "walk": wms, "table_update": float(upd_ms.mean())},
            "per_traversal": {"decisions": d / B, "update_nodes": u / B, "update_node_actions": v / B,
                              "leaves": float(tr["reserved"].astype(np.float64).mean())},
            "records_sorted": recs, "long_runs": long_runs, "pieces": pieces,
            "walk_roofline": {"bound": "hbm", "algorithmic_bytes_per_launch": walk_bytes,
                              "achieved": walk_bytes / (wms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": walk_bytes / (wms * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "tree_bytes_per_launch": tree_bytes, "tree_gbs": tree_bytes / (dt / steps) / 1e9}


def large_table_run(mccfr, table_log2, batch_log2, traversals, table_seed):
    """6 players, 200bb (4,000 chips), 10-action menu, <= 4 raises per street,
    2^28-slot tables (2^32 regret + strategy-sum entries, 32 GiB FP32, plus the
    8 GiB BF16 snapshot), launches of 2^22 traversals from the captured
    training graph; `traversals` in total (rounded to whole iterations)."""
    cfg = mccfr.cfg_6p_200bb_10(table_log2=table_log2, seed=0)
    B = 1 << batch_log2
    N = cfg.num_players
    iters = max(1, int(round(traversals / (B * N))))
    ctx = mccfr.Context(cfg, B)
    ctx.use_generated_hand_table()
    slots, row = ctx.slots, ctx.row
    log(f"bench: large_table 2^{cfg.table_log2} slots x {row} actions, batch 2^{B.bit_length() - 1}, "
        f"{iters} iterations x {N} launches")
    ctx.randomize_tables(table_seed)
    ctx.set_timing(True)
    for L in range(2 * N):  # warm-up, timed per kernel
        ctx.run_launch(L % N, L // N + 1, L * B, B)
    deal_ms, walk_ms, upd_ms = ctx.kernel_times(2 * N)
    ctx.set_timing(False)
    recs, long_runs, pieces = ctx.last_launch_stats()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.train(iters, B)
    dt = time.perf_counter() - t0
    total = iters * N * B
    result = {
        "metric": "MCCFR traversals/s (6-max 200bb, 10-action menu, 2^32-entry tables), 1 MI355X",
        "value": total / dt, "unit": "traversals/s", "n_gpus": 1, "steps": iters * N, "warmup": 2 * N,
        "ms_per_step": dt * 1e3 / (iters * N), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "large_table: 6 players, 4000-chip stacks, blinds 10/20, menu {F, C/C, 1/4, 1/2, "
                               "3/4, 1, 3/2, 2, 3 pot, AI}, <=4 raises/street",
                   "table_slots": slots, "row_actions": row, "entries": slots * row,
                   "table_bytes": slots * row * 10, "traversals_per_launch": B, "traversals": total,
                   "parallelism": "single GPU"},
        "kernel_ms": {"deal": float(deal_ms[N:].mean()), "walk": float(walk_ms[N:].mean()),
                      "table_update": float(upd_ms[N:].mean())},
        "launch": {"records_sorted": recs, "long_runs": long_runs, "pieces": pieces,
                   "table_atomics": 0,
                   # one atomic per walk wave on its emission block's record count, and one
                   # 64-bit counter atomic per long-run search wave (4,096 records) holding a long run
                   "counter_atomics_max": B // 64 + min(long_runs, -(-recs // 4096))},
    }
    for name in ("large_table_l2.json", "large_table_traffic.json"):
        path = os.path.join(ROOT, "profiles", name)
        if os.path.exists(path):
            with open(path) as f:
                result[name[len("large_table_"):-5]] = json.load(f)
    ctx.close()
    return result


def large_table_bench(args):
    import mccfr
    result = large_table_run(mccfr, args.table_log2 or 28, args.batch_log2 or 22, args.large_traversals,
                             args.table_seed)
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("6max_100bb", "large_table"), default="6max_100bb",
                    help="large_table: the 2^32-entry stress workload, one JSON line of its own")
    ap.add_argument("--large-traversals", type=float, default=1e9)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--batch-log2", type=int, default=None, help="default 20 (large_table: 22)")
    ap.add_argument("--table-log2", type=int, default=None, help="default 26 (large_table: 28)")
    ap.add_argument("--hand-log2", type=int, default=28)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0 = the process's host-core share (OMP_NUM_THREADS), capped by the visible CPUs")
    ap.add_argument("--cpu-launches", type=int, default=24)
    ap.add_argument("--batched-launches", type=int, default=8)
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-saturated", type=int, default=120)
    ap.add_argument("--serve-budget-gib", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-hand-bench", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-harness", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the reduced large-table leg")
    ap.add_argument("--no-es", action="store_true", help="skip the external-sampling leg")
    ap.add_argument("--large-leg-launches", type=int, default=12,
                    help="launches of 2^22 traversals in the default run's large-table leg")
    ap.add_argument("--profile", action="store_true", help="headline launches only (for rocprofv3 runs)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="with --profile: also run a plain device copy of this size (byte-counter calibration)")
    ap.add_argument("--gather-probe-mib", type=int, default=0,
                    help="with --profile: also run 2^24 random 16-byte reads over a table of this size")
    ap.add_argument("--profile-hands", type=int, default=0,
                    help="with --profile: also run the hand-evaluation microbenchmark on 2^N hands")
    ap.add_argument("--profile-es", type=int, default=0,
                    help="with --profile: also run this many external-sampling launches of 2^20 traversals "
                         "(after 6 warm-up launches, as the bench leg)")
    ap.add_argument("--walk-shape", default="",
                    help="with --profile: the walk's launch shape as WG:TPL (workgroup lanes : traversals "
                         "per lane, 0 = persistent waves); default the library's 256:1")
    ap.add_argument("--table-seed", type=int, default=2026,
                    help="seed of the synthetic regret state the timed launches start from")
    args = ap.parse_args()

    if args.walk_shape and not args.profile:
        sys.exit("bench.py: --walk-shape is for --profile runs (the timed line uses the default shape)")
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU delta exchange is design-only)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "large_table":
        large_table_bench(args)
        return
    args.batch_log2 = args.batch_log2 or 20
    args.table_log2 = args.table_log2 or 26

    import mccfr
    import workloads
    cfg = mccfr.cfg_6p_100bb(table_log2=args.table_log2, seed=0)
    B = 1 << args.batch_log2
    N = cfg.num_players
    log(f"bench: 6max_100bb table 2^{args.table_log2} batch 2^{args.batch_log2} warmup {args.warmup} "
        f"steps {args.steps}")
    ctx = mccfr.Context(cfg, B)
    ctx.use_generated_hand_table()
    if args.walk_shape:
        wg, tpl = (int(x) for x in args.walk_shape.split(":"))
        ctx.set_launch_shape(wg, tpl)
    hr = mccfr.generate_hand_table()
    state = {"launch": 0, "tid": 0}

    def step(trace=False):
        L = state["launch"]
        tr = ctx.run_launch(L % N, L // N + 1, state["tid"], B, trace=trace)
        state["launch"] += 1
        state["tid"] += B
        return tr

    fresh = None
    if not args.profile:
        # Fresh (all-zero) tables, as the reference's benchmark mode runs them
        # (main.cpp:380-398): every node takes the passive fallback, and the
        # 6-handed check-down ends at the depth guard.
        for _ in range(3):
            step()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(6):
            step()
        ctx.sync()
        fresh = 6 * B / (time.perf_counter() - t0)
        log(f"bench: fresh tables {fresh / 1e6:.1f} M traversals/s")

    # Timed state: seeded random regrets (mixed strategies at every info set),
    # so paths mix folds, calls, raises, all-ins and showdowns.
    ctx.randomize_tables(args.table_seed)
    t0 = time.perf_counter()
    for _ in range(args.warmup):
        step()
    ctx.sync()
    log(f"bench: warm-up done in {time.perf_counter() - t0:.2f} s")

    # `value`: the training loop, as the reference's benchmark mode times
    # train() (main.cpp:380-398): mccfr_train_tenants on the context's game,
    # `steps` launches (the update player cycling over the seats) replayed from
    # its captured hipGraph; one launch first captures the graph, untimed
    ctx.train_tenants([0], [B], 1)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.train_tenants([0], [B], args.steps)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    value = args.steps * B / elapsed
    # per-kernel times: the same number of launches issued one by one
    # (mccfr_run_launch, HIP events on the library stream around the phases)
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    loop_elapsed = time.perf_counter() - t0
    deal_ms, walk_ms, upd_ms = ctx.kernel_times(args.steps)
    ctx.set_timing(False)
    log(f"bench: {value / 1e6:.1f} M traversals/s (training graph; {args.steps * B / loop_elapsed / 1e6:.1f} M "
        f"launched one by one); deal {deal_ms.mean():.3f} ms, walk {walk_ms.mean():.3f} ms, "
        f"table update {upd_ms.mean():.3f} ms per launch")

    result = {
        "metric": METRIC, "value": value, "unit": "traversals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "6max_100bb: 6 players, 2000-chip stacks, blinds 10/20, menu "
                               "{F, C/C, 1/2P, 1P, 2P, AI}, <=3 raises/street",
                   "table_slots": 1 << args.table_log2, "traversals_per_launch": B, "parallelism": "single GPU"},
        "kernel_ms": {"deal": float(deal_ms.mean()), "walk": float(walk_ms.mean()),
                      "table_update": float(upd_ms.mean())},
        "table_state": f"seeded random regrets (seed {args.table_seed}) + {args.warmup} warm-up launches "
                       f"+ 1 launch capturing the training graph",
        "timed_as": "mccfr_train_tenants: the steps' launches replayed from the captured training graph",
        "launch_loop": {"traversals_per_s": args.steps * B / loop_elapsed,
                        "ms_per_step": loop_elapsed * 1e3 / args.steps,
                        "note": "the same launches issued one by one through mccfr_run_launch (host issue and "
                                "a segment upload per launch), where kernel_ms is timed"},
        "fresh_tables_traversals_per_s": fresh,
    }
    if args.profile:
        if args.copy_probe_mib:
            result["copy_probe"] = {"bytes": args.copy_probe_mib << 20,
                                    "ms": ctx.copy_probe(args.copy_probe_mib << 20, 2)}
        if args.gather_probe_mib:
            reads = 1 << 24
            result["gather_probe"] = {"table_bytes": args.gather_probe_mib << 20, "reads": reads,
                                      "bytes_read": 16 * reads, "bytes_written": 4 * reads,
                                      "ms": ctx.gather_probe(args.gather_probe_mib << 20, reads, 2)}
        if args.profile_hands:
            result["hand_eval"] = hand_bench(ctx, args.profile_hands)
        if args.profile_es:
            result["external_sampling"] = es_bench(mccfr, warmup=6, steps=args.profile_es,
                                                   table_seed=args.table_seed)
        print(json.dumps(result), flush=True)
        return

    tr = step(trace=True)  # path statistics of a launch like the timed ones
    recs, long_runs, pieces = ctx.last_launch_stats()
    wbytes = walk_bytes(tr, N, ctx.row)
    achieved = wbytes / (float(walk_ms.mean()) * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "walk_6max_100bb_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    result["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                          "kernel": "k_walk (betting-tree walk)",
                          "algorithmic_bytes_per_launch": wbytes,
                          "per_traversal": {"decisions": float(tr["decisions"].mean()),
                                            "update_nodes": float(tr["updates"].mean()),
                                            "update_legal_actions": float(tr["update_legal"].mean())}}
    # the walk's actual bound: vector issue, from the committed counter profile
    # (VALU instructions per launch) over this run's walk time
    ipath = os.path.join(ROOT, "profiles", "walk_6max_100bb_issue.json")
    if os.path.exists(ipath):
        with open(ipath) as f:
            iss = json.load(f)
        slots = float(walk_ms.mean()) * 1e-3 * iss["clock_hz"] * iss["simds"] / iss["cycles_per_instruction"]
        # `bound`/`frac` stay the HBM roofline the path is priced against;
        # what limits the walk in fact is vector instruction issue
        result["roofline"]["limiter"] = {"kind": "vector-issue",
                                         "valu_instructions_per_launch": iss["valu_instructions_per_launch"],
                                         "issue_frac": iss["valu_instructions_per_launch"] / slots,
                                         "lanes_active_per_instruction": iss["lanes_active_per_instruction"],
                                         "source": iss["source"]}
    result["launch"] = {"records_sorted": recs, "long_runs": long_runs, "pieces": pieces,
                        "path_bytes_per_launch": path_bytes(tr, N, ctx.row),
                        "path_gbs": path_bytes(tr, N, ctx.row) / (elapsed / args.steps) / 1e9}

    if not args.no_hand_bench:
        log("bench: hand-evaluation microbenchmark")
        result["hand_eval"] = hand_bench(ctx, args.hand_log2)

    threads, visible = host_cpus()
    if args.cpu_threads:
        threads = args.cpu_threads
    if not args.no_cpu_baseline:
        log(f"bench: CPU baseline on {threads} threads and on 1 ({visible} CPUs visible)")
        cb = cpu_baseline(mccfr, ctx, cfg, state["launch"], state["tid"], B, args.cpu_launches, threads)
        if cb:
            result["cpu_baseline"] = {
                "value": cb["threads_all"]["traversals_per_s"], "unit": "traversals/s", "cores": threads,
                "kind": "port", "cpus_visible": visible,
                "sample": f"{cb['threads_all']['launches']} launch(es) x {B} traversals on {threads} threads, "
                          f"{cb['threads_1']['launches']} on 1 thread, of the same workload from the GPU's "
                          f"post-warm-up tables (FP32, action-major, full regret-matching and clamp sweeps)",
                "threads_all": cb["threads_all"], "threads_1": cb["threads_1"]}
        if cb and not args.no_hand_bench:
            result["cpu_baseline"]["hand_eval_cpu"] = hand_eval_cpu(mccfr, ctx, 24, threads)
    if not args.no_harness:
        log("bench: strategy-quality harness")
        result["strategy_harness"] = harness_bench(mccfr, ctx, cfg)
        log(f"bench: harness {result['strategy_harness']['gpu_hands_per_s'] / 1e6:.1f} M hands/s on the GPU, "
            f"{result['strategy_harness']['cpu_reference_stream']['hands_per_s'] / 1e6:.3f} M hands/s "
            f"CPU (1 thread)")
    ctx.close()

    if not args.no_batched:
        log("bench: batched uneven games")
        result["batched"] = batched_bench(mccfr, workloads, hr, args.batched_launches)
        log(f"bench: batched {result['batched']['batched_traversals_per_s'] / 1e6:.1f} M/s vs plain "
            f"{result['batched']['plain_traversals_per_s'] / 1e6:.1f} M/s")

    if not args.no_cpu_baseline:
        log("bench: config 1 (3 players, 20bb) on the CPU and the GPU")
        result["config1"] = config1_bench(mccfr, hr, threads)

    if not args.no_es:
        log("bench: external sampling")
        result["external_sampling"] = es_bench(mccfr)
        log(f"bench: external sampling {result['external_sampling']['traversals_per_s'] / 1e6:.1f} M traversals/s "
            f"({result['external_sampling']['per_traversal']['update_nodes']:.1f} update nodes per traversal)")

    if not args.no_large:
        log("bench: large table (reduced run)")
        lt = large_table_run(mccfr, 28, 22, args.large_leg_launches * (1 << 22), args.table_seed)
        result["large_table"] = {k: lt[k] for k in ("metric", "value", "unit", "steps", "ms_per_step", "config",
                                                    "kernel_ms", "launch")}

    if not args.no_serving:
        log("bench: re-solve serving")
        sv = serving_bench(mccfr, workloads, hr, args.serve_requests, args.serve_saturated,
                           args.serve_budget_gib << 30, 256)
        result["serving"] = sv
        result["p99_resolve_latency_ms"] = sv["p99_ms"]
        log(f"bench: serving p50 {sv['p50_ms']:.1f} ms p99 {sv['p99_ms']:.1f} ms; saturated speedup "
            f"{sv['saturated']['speedup']:.2f}x over one at a time, "
            f"{sv['saturated']['speedup_vs_per_request_graph']:.2f}x over per-request graphs")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
