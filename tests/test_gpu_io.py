"""Table IO on the device: the average strategy, the .bin v2 strategy export
(poker_cuda/src/strategy.cpp:35-52 over get_strategy, cfr_gpu.cu:1383-1412), the
.ckpt checkpoint (cfr_gpu.cu:1464-1538) and the C++ drop-in headers run end to end.

Expected values come from the downloaded tables with the reference's float32
arithmetic restated in numpy: sums added left to right in action order,
S / total where total > 1e-7, else 0 (cfr_gpu.cu:1049-1057), rows kept where
the normalised row adds to >= 1e-6 (cfr_gpu.cu:1398-1401).
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import mccfr

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
V2, CKPT = 0x53545241540100, 0x43465247505500


def trained_ctx(hand_table, N=3, log2=14, iters=6, batch=4096, seed=5):
    cfg = mccfr.reference_cfg(num_players=N, starting_stack=400, table_log2=log2, seed=seed)
    ctx = mccfr.Context(cfg, batch)
    ctx.set_hand_table(hand_table)
    ctx.train(iters, batch)
    return ctx


def expected_normalized(S, row):
    S = S.reshape(-1, row)
    total = mccfr.row_totals(S)
    out = np.zeros_like(S)
    ok = total > np.float32(1e-7)
    out[ok] = (S[ok] / total[ok, None]).astype(np.float32)
    return out


def test_normalized_strategy_is_bit_exact(hand_table):
    with trained_ctx(hand_table) as ctx:
        _, S, _ = ctx.download()
        P = ctx.normalized_strategy().reshape(-1, ctx.row)
        E = expected_normalized(S, ctx.row)
        assert np.count_nonzero(E.sum(1)) > 100
        assert np.array_equal(P.view(np.uint32), E.view(np.uint32))
        # action-major layout is the transpose
        Pt = ctx.normalized_strategy(layout=1).reshape(ctx.row, -1)
        assert np.array_equal(Pt.T, P)


def test_save_strategy_writes_get_strategy_rows(hand_table, tmp_path):
    with trained_ctx(hand_table) as ctx:
        _, S, _ = ctx.download()
        E = expected_normalized(S, ctx.row)
        keep = np.nonzero(mccfr.row_totals(E) >= np.float32(1e-6))[0].astype(np.uint32)
        path = str(tmp_path / "s.bin")
        ctx.save_strategy(path)
        raw = open(path, "rb").read()
        magic, ts, na, n = struct.unpack_from("<QIII", raw)
        assert (magic, ts, na, n) == (V2, ctx.slots, ctx.row, keep.size)
        rec = np.frombuffer(raw, dtype=np.dtype([("k", "<u4"), ("p", "<f4", (ctx.row,))]), offset=20)
        assert np.array_equal(rec["k"], keep)
        assert np.array_equal(rec["p"].view(np.uint32), E[keep].view(np.uint32))
        # the loader gives the same rows back
        ts2, na2, keys, probs = mccfr.load_strategy(path)
        assert (ts2, na2) == (ctx.slots, ctx.row) and np.array_equal(keys, keep)
        # the Python trainer's get_strategy picks the same rows
        tr = mccfr.Trainer.__new__(mccfr.Trainer)
        tr.ctx = ctx
        got = tr.get_strategy()
        assert sorted(got) == keep.tolist()


def test_checkpoint_layout_and_resume(hand_table, tmp_path):
    """The .ckpt holds u64 magic, u32 table_size, u32 num_actions, then regrets and
    strategy sums action-major (cfr_gpu.cu:1464-1491); a context loaded from it
    continues exactly as the one that wrote it."""
    path = str(tmp_path / "c.ckpt")
    with trained_ctx(hand_table, iters=4) as a:
        a.save_checkpoint(path)
        R, S, _ = a.download(layout=1)
        raw = open(path, "rb").read()
        assert struct.unpack_from("<QII", raw) == (CKPT, a.slots, a.row)
        body = np.frombuffer(raw, np.float32, offset=16)
        assert body.size == 2 * R.size
        assert np.array_equal(body[:R.size].view(np.uint32), R.view(np.uint32))
        assert np.array_equal(body[R.size:].view(np.uint32), S.view(np.uint32))
        cfg = a.cfg
        with mccfr.Context(cfg, 4096) as b:
            b.set_hand_table(hand_table)
            b.load_checkpoint(path)
            Ra, Sa, siga = a.download()
            Rb, Sb, sigb = b.download()
            assert np.array_equal(Ra.view(np.uint32), Rb.view(np.uint32))
            assert np.array_equal(Sa.view(np.uint32), Sb.view(np.uint32))
            assert np.array_equal(siga, sigb)
            # same launch on both: same tables (traversal ids and weights given explicitly)
            for ctx in (a, b):
                ctx.run_launch(1, 5, 10_000_000, 4096)
            Ra, Sa, _ = a.download()
            Rb, Sb, _ = b.download()
            assert np.array_equal(Ra.view(np.uint32), Rb.view(np.uint32))
            assert np.array_equal(Sa.view(np.uint32), Sb.view(np.uint32))
        # a strategy file is not a checkpoint
        a.save_strategy(str(tmp_path / "s.bin"))
        with pytest.raises(mccfr.MccfrError) as ei:
            a.load_checkpoint(str(tmp_path / "s.bin"))
        assert ei.value.code == mccfr.E_FORMAT


def test_cpp_headers_run_end_to_end(tmp_path):
    """GPUCFRTrainer + strategy_save / strategy_load / strategy_load_auto at the
    reference geometry (2^26 slots x 8 actions), as the reference CLI calls them."""
    src = os.path.join(ROOT, "tests", "cpp", "wrapper_usage.cpp")
    exe = tmp_path / "wrapper_usage_exe"
    pkg = os.path.dirname(mccfr.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), src, "-L", pkg, "-lmccfr",
                    "-Wl,-rpath," + pkg, "-o", str(exe)], check=True)
    print("wrapper_usage: 2 iterations x 4096, 2 players, 2^26 slots", flush=True)
    r = subprocess.run([str(exe), "-", str(tmp_path)], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "Exported playable strategy to" in r.stderr
    assert os.path.exists(tmp_path / "wrapper_usage")
