"""Strategy files in the reference's .bin v2 layout (poker_cuda/src/strategy.cpp:6-156).

The byte layout is pinned against a file assembled here from the layout the
reference documents (strategy.cpp:6-8: u64 magic, u32 table_size, u32
num_actions, u32 entries, then u32 key + f32 probs[8] per entry) with the
reference's magic values (strategy.cpp:10-12).  These tests make no device call.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import mccfr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
V1, V2, CKPT = 0x53545241540000, 0x53545241540100, 0x43465247505500


def assemble(magic, table_size, na, keys, probs):
    b = struct.pack("<QIII", magic, table_size, na, len(keys))
    for k, row in zip(keys, probs):
        b += struct.pack("<I", int(k)) + struct.pack(f"<{na}f", *[float(x) for x in row])
    return b


def rows(n, na=8, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.random((n, na)).astype(np.float32)
    return (p / p.sum(1, keepdims=True)).astype(np.float32)


def test_writer_bytes_match_the_documented_layout(tmp_path):
    keys = np.array([0, 5, 77, (1 << 26) - 1], np.uint32)
    probs = rows(4)
    path = str(tmp_path / "s.bin")
    mccfr.write_strategy(path, 1 << 26, keys, probs)
    assert open(path, "rb").read() == assemble(V2, 1 << 26, 8, keys, probs)


def test_round_trip_is_bit_exact(tmp_path):
    keys = np.sort(np.random.default_rng(1).choice(1 << 20, 500, replace=False)).astype(np.uint32)
    probs = rows(500, seed=2)
    path = str(tmp_path / "s.bin")
    mccfr.write_strategy(path, 1 << 20, keys, probs)
    ts, na, k2, p2 = mccfr.load_strategy(path)
    assert (ts, na) == (1 << 20, 8)
    assert np.array_equal(k2, keys)
    assert np.array_equal(p2.view(np.uint32), probs.view(np.uint32))


def test_empty_file_and_wide_rows(tmp_path):
    path = str(tmp_path / "e.bin")
    mccfr.write_strategy(path, 1 << 14, np.zeros(0, np.uint32), np.zeros((0, 8), np.float32))
    assert open(path, "rb").read() == assemble(V2, 1 << 14, 8, [], [])
    ts, na, k, p = mccfr.load_strategy(path)
    assert (ts, na, k.size, p.shape) == (1 << 14, 8, 0, (0, 8))
    # a 10-action menu writes 10 probabilities per row
    keys, probs = np.array([3, 9], np.uint32), rows(2, na=10)
    mccfr.write_strategy(path, 1 << 14, keys, probs)
    assert open(path, "rb").read() == assemble(V2, 1 << 14, 10, keys, probs)
    assert np.array_equal(mccfr.load_strategy(path)[3], probs)


def test_uniform_placeholder_rows_are_skipped(tmp_path):
    """strategy.cpp:22-33: rows within 1e-6 of 1/8 in every action are dropped."""
    u = np.full((4, 8), 0.125, np.float32)
    u[1, 3] += 5e-7          # still a placeholder
    u[2, 0] += 2e-6          # a real row (one action beyond 1e-6)
    u[3] = rows(1)[0]        # a real row
    path = tmp_path / "p.bin"
    path.write_bytes(assemble(V2, 1 << 26, 8, [10, 11, 12, 13], u))
    ts, na, k, p = mccfr.load_strategy(str(path))
    assert k.tolist() == [12, 13]
    assert np.array_equal(p, u[2:])


@pytest.mark.parametrize("magic", [CKPT, V1, 0x1234])
def test_foreign_headers_are_format_errors(tmp_path, magic):
    path = tmp_path / "x.bin"
    path.write_bytes(assemble(magic, 1 << 26, 8, [1], rows(1)))
    with pytest.raises(mccfr.MccfrError) as ei:
        mccfr.load_strategy(str(path))
    assert ei.value.code == mccfr.E_FORMAT


def test_truncated_and_missing_files(tmp_path):
    path = tmp_path / "t.bin"
    path.write_bytes(assemble(V2, 1 << 26, 8, [1, 2, 3], rows(3))[:-5])
    with pytest.raises(mccfr.MccfrError) as ei:
        mccfr.load_strategy(str(path))
    assert ei.value.code == mccfr.E_FORMAT
    with pytest.raises(mccfr.MccfrError) as ei:
        mccfr.load_strategy(str(tmp_path / "missing.bin"))
    assert ei.value.code == mccfr.E_IO


def test_cpp_strategy_functions(tmp_path):
    """include/mccfr/strategy.hpp (strategy_save / load / load_auto / print_top) on the host."""
    src = os.path.join(ROOT, "tests", "cpp", "strategy_io.cpp")
    exe = tmp_path / "strategy_io"
    pkg = os.path.dirname(mccfr.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src, "-L", pkg, "-lmccfr",
                    "-Wl,-rpath," + pkg, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    # strategy.cpp:158-181 prints "  %-16s %.1f%%" for the top actions above 0.1 %
    assert r.stdout.splitlines() == ["  raise_pot        50.0%", "  call             30.0%", "  fold             20.0%",
                                     "(info set not found)"]
    assert "is a training checkpoint" in r.stderr
    assert "older (v1)" in r.stderr
    assert "expects 67108864 and 8" in r.stderr
