"""The C-ABI library loads and exports every symbol its header declares; host-only
entry points work without a device."""
import ctypes as C
import os
import re

import numpy as np

import mccfr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mccfr", "mccfr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mccfr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(mccfr.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(mccfr.exported_symbols())


def test_generator_and_preflop_lut_are_host_code(golden_lut):
    hr = mccfr.generate_hand_table()
    assert hr.size == mccfr.HAND_TABLE_WORDS
    assert np.array_equal(mccfr.preflop_lut(), golden_lut)


def test_reference_cfg_defaults():
    cfg = mccfr.reference_cfg()
    assert (cfg.num_players, cfg.starting_stack, cfg.small_blind, cfg.big_blind) == (2, 1000, 10, 20)
    assert (cfg.num_actions, cfg.history_bits, cfg.table_log2, cfg.max_raises) == (8, 3, 26, 0)
    kinds = [cfg.menu[i].kind for i in range(8)]
    assert kinds == [mccfr.FOLD, mccfr.CHECK, mccfr.CALL, mccfr.RAISE_MIN, mccfr.POT_FRAC, mccfr.POT_FRAC,
                     mccfr.POT_FRAC, mccfr.ALL_IN]
    assert [(cfg.menu[i].num, cfg.menu[i].den) for i in (4, 5, 6)] == [(1, 2), (2, 3), (1, 1)]


def test_bad_arguments_are_errors_not_crashes():
    L = mccfr.lib()
    assert L.mccfr_generate_hand_table(None, 10, 0) == -1
    h = C.c_void_p()
    cfg = mccfr.reference_cfg()
    cfg.num_players = 9
    assert L.mccfr_create(C.byref(cfg), 1024, C.byref(h)) == -1
    assert b"num_players" in L.mccfr_last_error(None)
    cfg = mccfr.reference_cfg()
    assert L.mccfr_create(C.byref(cfg), 0, C.byref(h)) == -1


def test_cpp_wrapper_compiles_and_links(tmp_path):
    """include/mccfr/gpu_cfr_trainer.hpp builds against the reference CLI's call pattern."""
    import subprocess
    src = os.path.join(ROOT, "tests", "cpp", "wrapper_usage.cpp")
    out = tmp_path / "wrapper_usage"
    pkg = os.path.dirname(mccfr.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src, "-L", pkg, "-lmccfr",
                    "-Wl,-rpath," + pkg, "-o", str(out)], check=True)
    assert out.exists()
