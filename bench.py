#!/usr/bin/env python3
"""Benchmark of the MCCFR hot path on one MI355X.

Workload (default): 6-max 100bb NLH -- 6 players, 2,000-chip stacks at
blinds 10/20, menu {fold, check/call, 1/2 pot, pot, 2 pot, all-in}, at most 3
raises per street; FP32 regret and strategy-sum tables of 2^26 slots; one step =
one launch of 2^20 traversals (deal, buckets, betting-tree walk, delta scatter,
touched-slot CFR+ update), update player cycling over the seats as in the
training loop.  Inputs (tables, hand table) are resident in HBM before timing.

Prints one JSON line with traversals/s, the traversal kernel's HBM roofline
fraction (HIP events on the library stream around every timed launch), the
hand-evaluation microbenchmark (2^28 hands, both evaluators) and the CPU
baseline (the oracle's FP32 restatement of the reference launch, timed on a
bounded sample on this host).
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


def traversal_bytes(tr, num_players, row):
    """Algorithmic bytes of a launch (per-traversal figure x traversals):
    72*N hand-table words (5+6+7 words x 4 B per player), 2*A B per decision
    node (BF16 strategy row), 4*A B per update node (FP32 regret row) and 16 B
    per legal action at an update node (read-modify-write of regret and
    strategy sum, 4+4 B each)."""
    d = tr["decisions"].astype(np.float64).sum()
    u = tr["updates"].astype(np.float64).sum()
    v = tr["update_legal"].astype(np.float64).sum()
    return 72.0 * num_players * len(tr) + 2.0 * row * d + 4.0 * row * u + 16.0 * v


def hand_bench(mccfr, ctx, log2n, reps=3):
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
            out[name] = {"hands_per_s": n / dt, "ms": dt * 1e3,
                         "gbs_37B_per_hand": 37.0 * n / dt / 1e9}
    finally:
        ctx.dfree(dc)
        ctx.dfree(dr)
    out["hands"] = n
    return out


def cpu_baseline(mccfr, ctx, cfg, launch, tid, batch, launches, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    O = oracle.lib()
    R, S, _ = ctx.download(layout=1)  # same table state as the GPU, action-major like the reference
    sig = np.zeros(R.size, np.uint16)
    hr = mccfr.generate_hand_table()
    lut = mccfr.preflop_lut()
    N = cfg.num_players
    secs = O.orc_cpu_baseline(C.byref(cfg), oracle.p(hr), oracle.p(lut), oracle.p(R), oracle.p(S), oracle.p(sig),
                              launches, batch, launch % N, launch // N + 1, tid, threads)
    if secs <= 0:
        return None
    return {"value": launches * batch / secs, "unit": "traversals/s", "cores": threads, "kind": "port",
            "sample": f"{launches} launch(es) x {batch} traversals of the same workload from the GPU's "
                      f"post-warm-up tables (FP32, action-major, full regret-matching and clamp sweeps), "
                      f"{secs:.1f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--batch-log2", type=int, default=20)
    ap.add_argument("--table-log2", type=int, default=26)
    ap.add_argument("--hand-log2", type=int, default=28)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, host cpus)")
    ap.add_argument("--cpu-launches", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-hand-bench", action="store_true")
    ap.add_argument("--profile", action="store_true", help="timed launches only (for rocprofv3 runs)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="with --profile: also run a plain device copy of this size (byte-counter calibration)")
    ap.add_argument("--table-seed", type=int, default=2026,
                    help="seed of the synthetic regret state the timed launches start from")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU delta exchange is design-only)", file=sys.stderr)
        sys.exit(2)

    import mccfr
    cfg = mccfr.cfg_6p_100bb(table_log2=args.table_log2, seed=0)
    B = 1 << args.batch_log2
    N = cfg.num_players
    log(f"bench: 6max_100bb table 2^{args.table_log2} batch 2^{args.batch_log2} warmup {args.warmup} "
        f"steps {args.steps}")
    ctx = mccfr.Context(cfg, B)
    ctx.use_generated_hand_table()
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

    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    elapsed = time.perf_counter() - t0
    trav_ms, upd_ms = ctx.kernel_times(args.steps)
    ctx.set_timing(False)
    value = args.steps * B / elapsed
    log(f"bench: {value / 1e6:.1f} M traversals/s; traverse kernel {trav_ms.mean():.3f} ms, "
        f"update kernel {upd_ms.mean():.3f} ms per launch")

    result = {
        "metric": METRIC, "value": value, "unit": "traversals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "6max_100bb: 6 players, 2000-chip stacks, blinds 10/20, menu "
                               "{F, C/C, 1/2P, 1P, 2P, AI}, <=3 raises/street",
                   "table_slots": 1 << args.table_log2, "traversals_per_launch": B, "parallelism": "single GPU"},
        "kernel_ms": {"traverse": float(trav_ms.mean()), "update": float(upd_ms.mean())},
        "table_state": f"seeded random regrets (seed {args.table_seed}) + {args.warmup} warm-up launches",
        "fresh_tables_traversals_per_s": fresh,
    }
    if args.profile:
        if args.copy_probe_mib:
            result["copy_probe"] = {"bytes": args.copy_probe_mib << 20,
                                    "ms": ctx.copy_probe(args.copy_probe_mib << 20, 2)}
        print(json.dumps(result), flush=True)
        return

    tr = step(trace=True)  # path statistics of a launch like the timed ones
    nbytes = traversal_bytes(tr, N, ctx.row)
    achieved = nbytes / (float(trav_ms.mean()) * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traverse_6max_100bb_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    result["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                          "algorithmic_bytes_per_launch": nbytes,
                          "per_traversal": {"decisions": float(tr["decisions"].mean()),
                                            "update_nodes": float(tr["updates"].mean()),
                                            "update_legal_actions": float(tr["update_legal"].mean())}}

    if not args.no_hand_bench:
        log("bench: hand-evaluation microbenchmark")
        result["hand_eval"] = hand_bench(mccfr, ctx, args.hand_log2)

    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"bench: CPU baseline on {threads} threads")
        result["cpu_baseline"] = cpu_baseline(mccfr, ctx, cfg, state["launch"], state["tid"], B,
                                              args.cpu_launches, threads)
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
