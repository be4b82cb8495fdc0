"""poker_mccfr, the command-line trainer (reference CLI poker_cuda/src/main.cpp:29-69,
:98-139, :295-398): flags, defaults and argument errors on the host; train,
resume and benchmark runs on the device.

A CLI run is checked against the same training done through the Python
bindings: the .ckpt and .bin files it writes must be byte-identical to the
ones save_checkpoint / save_strategy write for that game.
"""
import os
import re
import struct
import subprocess

import pytest

import mccfr

EXE = os.path.join(os.path.dirname(mccfr.LIB_PATH), "poker_mccfr")


def run(*args, timeout=120, cwd=None):
    return subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=timeout, cwd=cwd)


def test_help_lists_the_reference_flags():
    r = run("--help")
    assert r.returncode == 0
    for flag, default in [("--mode", "train"), ("--players", "2"), ("--iters", "1000"), ("--batch", "65536"),
                          ("--stack", "1000"), ("--sb", "10"), ("--bb", "20"), ("--save", "strategy.bin"),
                          ("--handranks", "data/handranks.dat"), ("--seed", "42"), ("--table-log2", "26"),
                          ("--raise-cap", "0")]:
        line = next(l for l in r.stdout.splitlines() if l.strip().startswith(flag + " "))
        assert f"default: {default}" in line, line
    for flag in ("--load", "--no-cfrplus", "--cfrplus", "--no-lcfr", "--lcfr", "--menu", "--batch-sizes",
                 "--bench-iters", "--strategy", "--opponent", "--hands"):
        assert flag in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["--menu", "fold,call,3/4,bogus"], "unknown --menu action 'bogus'"),
    (["--menu", "fold,0/2"], "unknown --menu action '0/2'"),
    (["--menu", ",".join(["pot"] * 17)], "--menu needs 1..16 actions"),
    (["--players", "9"], "num_players"),
    (["--table-log2", "40"], "table_log2"),
    (["--batch-sizes", "4096,x"], "bad --batch-sizes entry 'x'"),
    (["--mode", "play"], "not part of poker_mccfr"),
    (["--mode", "eval", "--players", "10"], "between 2 and 9 for eval mode"),
    (["--mode", "fly"], "unknown --mode fly"),
    (["--frobnicate"], "unknown or incomplete option --frobnicate"),
    (["--iters"], "unknown or incomplete option --iters"),
])
def test_argument_errors_exit_2_before_device_work(args, msg):
    r = run(*args)
    assert r.returncode == 2, (r.stdout, r.stderr)
    assert msg in r.stderr
    assert "poker_mccfr (MI355X)" not in r.stdout


def cli_game(players, table_log2, seed, menu=None, raise_cap=0):
    if menu is None:
        cfg = mccfr.reference_cfg(num_players=players, table_log2=table_log2, seed=seed)
    else:
        cfg = mccfr.menu_cfg(players, 1000, menu, max_raises=raise_cap, table_log2=table_log2, seed=seed)
    cfg.max_raises = raise_cap
    return cfg


def python_files(cfg, iters, batch, out, load=None):
    with mccfr.Context(cfg, batch) as ctx:
        ctx.use_generated_hand_table()
        if load:
            ctx.load_checkpoint(load)
        ctx.train(iters, batch)
        ctx.save_checkpoint(out + ".ckpt")
        ctx.save_strategy(out)


@pytest.mark.gpu
@pytest.mark.parametrize("menu,cap", [(None, 0), ("fold,checkcall,1/2,pot,allin", 2)])
def test_train_writes_the_same_files_as_the_library(tmp_path, menu, cap):
    s = tmp_path / "cli.bin"
    args = ["--players", 3, "--iters", 3, "--batch", 4096, "--table-log2", 16, "--seed", 7, "--save", s,
            "--handranks", "generated", "--raise-cap", cap]
    if menu:
        args += ["--menu", menu]
    print("poker_mccfr train", args, flush=True)
    r = run(*args)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    assert "Checkpoint saved to" in r.stdout and "Strategy saved to" in r.stdout
    kinds = {"fold": mccfr.FOLD, "checkcall": mccfr.CHECK_CALL, "pot": mccfr.POT_FRAC, "allin": mccfr.ALL_IN}
    m = None
    if menu:
        m = [(mccfr.POT_FRAC, 1, 2) if t == "1/2" else (mccfr.POT_FRAC, 1, 1) if t == "pot" else (kinds[t], 0, 1)
             for t in menu.split(",")]
    cfg = cli_game(3, 16, 7, m, cap)
    p = str(tmp_path / "py.bin")
    python_files(cfg, 3, 4096, p)
    assert open(str(s) + ".ckpt", "rb").read() == open(p + ".ckpt", "rb").read()
    assert open(s, "rb").read() == open(p, "rb").read()
    n = int(re.search(r"\((\d+) info sets\)", r.stdout).group(1))
    assert n > 0 and n == struct.unpack_from("<I", open(s, "rb").read(), 16)[0]


@pytest.mark.gpu
def test_resume_from_checkpoint(tmp_path):
    a = tmp_path / "a.bin"
    common = ["--players", 2, "--batch", 4096, "--table-log2", 15, "--handranks", "generated"]
    print("poker_mccfr train + resume", flush=True)
    assert run(*common, "--iters", 2, "--save", a).returncode == 0
    b = tmp_path / "b.bin"
    r = run(*common, "--iters", 2, "--save", b, "--load", str(a) + ".ckpt")
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Loading checkpoint from" in r.stdout
    cfg = cli_game(2, 15, 42)
    p = str(tmp_path / "py.bin")
    python_files(cfg, 2, 4096, p, load=str(a) + ".ckpt")
    assert open(str(b) + ".ckpt", "rb").read() == open(p + ".ckpt", "rb").read()
    assert open(b, "rb").read() == open(p, "rb").read()


@pytest.mark.gpu
def test_benchmark_mode_prints_one_line_per_batch_size():
    print("poker_mccfr benchmark", flush=True)
    r = run("--mode", "benchmark", "--players", 6, "--table-log2", 20, "--batch-sizes", "4096,65536",
            "--bench-iters", 3, "--handranks", "generated")
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip().startswith("batch=")]
    assert len(lines) == 2
    for bs, l in zip((4096, 65536), lines):
        m = re.match(r"\s+batch=\s*(\d+)\s+hands=(\S+)M\s+time=(\S+)s\s+speed=(\S+)M/s", l)
        assert m and int(m.group(1)) == bs, l
        assert float(m.group(4)) > 0
    assert "setup excluded" in r.stdout


def eval_lines(out):
    vals = {}
    for key, pat in [("hands", r"Hands played : (\d+)"), ("net_bb", r"Net BB       : (\S+)"),
                     ("bb100", r"BB/100       : (\S+)"), ("loaded", r"Loaded (\d+) info sets")]:
        m = re.search(pat, out)
        vals[key] = float(m.group(1)) if m else None
    return vals


@pytest.mark.gpu
def test_eval_mode_on_strategy_and_checkpoint(tmp_path, hand_table):
    """main.cpp:146-198 on the device: a .bin is played from its rows, a .ckpt is
    converted through the trainer (and the converted .bin written next to it);
    both give the library's result for the same hands."""
    s = tmp_path / "e.bin"
    common = ["--players", 3, "--table-log2", 15, "--handranks", "generated"]
    assert run(*common, "--iters", 4, "--batch", 4096, "--save", s).returncode == 0
    print("poker_mccfr eval", flush=True)
    r = run("--mode", "eval", "--strategy", s, "--opponent", "nit", "--hands", 20000, *common)
    assert r.returncode == 0, r.stderr[-2000:]
    got = eval_lines(r.stdout)
    cfg = cli_game(3, 15, 42)
    ts, na, keys, probs = mccfr.load_strategy(str(s))
    with mccfr.Context(cfg, 4096) as ctx:
        ctx.use_generated_hand_table()
        want = ctx.evaluate_rows(cfg, keys, probs, 3, mccfr.NIT, 20000, seed=42)
    assert got["hands"] == 20000 and got["loaded"] == keys.size
    assert got["net_bb"] == pytest.approx(want.net_bb, abs=0.051)
    assert got["bb100"] == pytest.approx(want.bb_per_100, abs=0.0051)
    # the checkpoint gives the same hands: its normalised rows are the .bin's rows
    ck = tmp_path / "e.bin.ckpt"
    before = open(s, "rb").read()
    os.remove(s)
    r2 = run("--mode", "eval", "--strategy", ck, "--opponent", "nit", "--hands", 20000, *common)
    assert r2.returncode == 0, r2.stderr[-2000:]
    assert "Exported playable strategy to" in r2.stderr
    assert eval_lines(r2.stdout)["bb100"] == got["bb100"]
    # x.ckpt converts to x (strategy.cpp:140-146): here the .bin training wrote
    assert open(s, "rb").read() == before
    # nine seats against random opponents run too
    r3 = run("--mode", "eval", "--strategy", s, "--opponent", "random", "--hands", 5000, "--players", 9,
             "--table-log2", 15, "--handranks", "generated")
    assert r3.returncode == 0, r3.stderr[-2000:]
    assert "=== Eval Results (9-player) ===" in r3.stdout


@pytest.mark.gpu
def test_train_from_the_build_time_table_file(tmp_path):
    """The file gen_table writes (the build's data/handranks.dat) trains
    exactly as the in-memory table does: same checkpoint and strategy bytes;
    the default --handranks path is relative to the working directory, as in
    the reference (main.cpp)."""
    gen = os.path.join(os.path.dirname(mccfr.LIB_PATH), "gen_table")
    data = tmp_path / "data"
    data.mkdir()
    r = subprocess.run([gen, str(data / "handranks.dat")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    common = ["--players", 3, "--batch", 4096, "--table-log2", 15, "--iters", 2]
    print("poker_mccfr train from the table file and from the generated table", flush=True)
    a = tmp_path / "file.bin"
    r = run(*common, "--save", a, cwd=tmp_path)  # default: data/handranks.dat
    assert r.returncode == 0, r.stderr[-2000:]
    assert "NOTE:" not in r.stderr  # the file was found, no fallback
    b = tmp_path / "gen.bin"
    r = run(*common, "--save", b, "--handranks", "generated", cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    assert open(a, "rb").read() == open(b, "rb").read()
    assert open(str(a) + ".ckpt", "rb").read() == open(str(b) + ".ckpt", "rb").read()
