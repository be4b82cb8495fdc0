"""ctypes binding of libmccfr.so (the C ABI in include/mccfr/mccfr.h).

Python-side mirror of the reference trainer interface for tests and the
benchmark.  The library is the built ``libmccfr.so`` next to this file; a
missing library raises ``MccfrError`` -- there is no CPU fallback.

``Trainer`` follows ``GPUCFRTrainer`` (poker_cuda/src/cfr_gpu.cuh:48-132):
constructor arguments, ``load_hand_table``, ``train``, ``get_strategy``,
``export_tables``/``import_tables``, ``num_info_sets_active``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmccfr.so")

MAX_ACTIONS = 16
HAND_TABLE_WORDS = 32487834

FOLD, CHECK, CALL, CHECK_CALL, RAISE_MIN, POT_FRAC, ALL_IN = range(7)
SAMPLE_OUTCOME, SAMPLE_EXTERNAL = 0, 1  # mccfr_game_cfg.sampling


# scripted archetypes of the strategy harness (MCCFR_OPP_*, eval.h:5-22)
CALLING_STATION, NIT, MANIAC, BALANCED, RANDOM = range(5)


class EvalResult(C.Structure):
    _fields_ = [("hands_played", C.c_int64), ("net_chips", C.c_int64), ("net_bb", C.c_double),
                ("bb_per_100", C.c_double)]


# error codes of the C ABI (MCCFR_E_*)
E_ARG, E_HIP, E_TABLE, E_IO, E_FORMAT, E_NOMEM, E_STATE = -1, -2, -3, -4, -5, -6, -7


class MccfrError(RuntimeError):
    def __init__(self, msg: str, code: Optional[int] = None):
        super().__init__(msg)
        self.code = code


class Action(C.Structure):
    _fields_ = [("kind", C.c_uint8), ("num", C.c_uint8), ("den", C.c_uint8), ("reserved", C.c_uint8)]


class GameCfg(C.Structure):
    _fields_ = [
        ("num_players", C.c_int32),
        ("starting_stack", C.c_int32),
        ("small_blind", C.c_int32),
        ("big_blind", C.c_int32),
        ("use_cfr_plus", C.c_int32),
        ("use_linear_cfr", C.c_int32),
        ("num_actions", C.c_int32),
        ("max_raises", C.c_int32),
        ("history_bits", C.c_int32),
        ("table_log2", C.c_int32),
        ("seed", C.c_uint64),
        ("menu", Action * MAX_ACTIONS),
        ("sampling", C.c_int32),
        ("reserved", C.c_int32),
    ]


TRACE_DTYPE = np.dtype(
    [("utility", "<f4"), ("path_hash", "<u4"), ("decisions", "<u2"), ("updates", "u1"), ("terminal", "u1"),
     ("update_legal", "<u2"), ("reserved", "<u2")]
)


class RootState(C.Structure):
    """mccfr_root_state: street 0 = the reference's preflop root; 1/2/3 = a
    public flop/turn/river state (fixed board, stacks, investments, active seats)."""
    _fields_ = [
        ("street", C.c_int32),
        ("dealer", C.c_int32),
        ("active_mask", C.c_uint32),
        ("history", C.c_uint32),
        ("stack", C.c_int32 * 6),
        ("invested", C.c_int32 * 6),
        ("board", C.c_uint8 * 5),
        ("reserved", C.c_uint8 * 3),
    ]


def root_state(street, dealer, active, stacks, invested, board, history=0) -> RootState:
    r = RootState()
    r.street = street
    r.dealer = dealer
    r.active_mask = active if isinstance(active, int) else sum(1 << p for p in active)
    r.history = history
    for i, v in enumerate(stacks):
        r.stack[i] = int(v)
    for i, v in enumerate(invested):
        r.invested[i] = int(v)
    for i, c in enumerate(board):
        r.board[i] = int(c)
    return r


class Segment(C.Structure):
    _fields_ = [
        ("tenant", C.c_uint32),
        ("n", C.c_uint32),
        ("update_player", C.c_int32),
        ("reserved", C.c_uint32),
        ("t_weight", C.c_int64),
        ("first_tid", C.c_uint64),
    ]


class Trace(C.Structure):
    _fields_ = [
        ("utility", C.c_float),
        ("path_hash", C.c_uint32),
        ("decisions", C.c_uint16),
        ("updates", C.c_uint8),
        ("terminal", C.c_uint8),
        ("update_legal", C.c_uint16),
        ("reserved", C.c_uint16),
    ]



_lib: Optional[C.CDLL] = None

# name -> (restype, argtypes)
_SIGS = {
    "mccfr_last_error": (C.c_char_p, [C.c_void_p]),
    "mccfr_game_cfg_reference": (None, [C.POINTER(GameCfg), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mccfr_create": (C.c_int, [C.POINTER(GameCfg), C.c_uint32, C.POINTER(C.c_void_p)]),
    "mccfr_destroy": (None, [C.c_void_p]),
    "mccfr_generate_hand_table": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int]),
    "mccfr_set_hand_table": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "mccfr_load_hand_table": (C.c_int, [C.c_void_p, C.c_char_p]),
    "mccfr_use_generated_hand_table": (C.c_int, [C.c_void_p]),
    "mccfr_validate_hand_table": (C.c_int, [C.c_void_p, C.c_size_t]),
    "mccfr_validate_hand_table_file": (C.c_int, [C.c_char_p]),
    "mccfr_preflop_lut": (C.c_int, [C.c_void_p]),
    "mccfr_eval7_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]),
    "mccfr_eval7_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]),
    "mccfr_run_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_uint32, C.c_void_p]),
    "mccfr_run_launch_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_uint32]),
    "mccfr_train": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint32]),
    "mccfr_sync": (C.c_int, [C.c_void_p]),
    "mccfr_set_max_batch": (C.c_int, [C.c_void_p, C.c_uint32]),
    "mccfr_randomize_tables": (C.c_int, [C.c_void_p, C.c_uint64]),
    "mccfr_set_launch_shape": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "mccfr_get_launch_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "mccfr_save_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p]),
    "mccfr_load_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p]),
    "mccfr_table_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "mccfr_download_tables": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "mccfr_upload_tables": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "mccfr_normalized_strategy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "mccfr_counters": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_int64)]),
    "mccfr_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "mccfr_kernel_times": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "mccfr_copy_probe": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_float)]),
    "mccfr_gather_probe": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_int, C.POINTER(C.c_float)]),
    "mccfr_random_hands_device": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_size_t]),
    "mccfr_device_alloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "mccfr_device_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mccfr_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "mccfr_last_launch_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_uint32)]),
    "mccfr_save_strategy": (C.c_int, [C.c_void_p, C.c_uint32, C.c_char_p]),
    "mccfr_write_strategy": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "mccfr_load_strategy": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p, C.c_size_t]),
    "mccfr_evaluate_strategy": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_int64, C.c_uint64,
                                          C.c_void_p, C.c_void_p]),
    "mccfr_evaluate_rows": (C.c_int, [C.c_void_p, C.POINTER(GameCfg), C.c_void_p, C.c_void_p, C.c_uint32,
                                      C.c_uint32, C.c_int, C.c_int, C.c_int64, C.c_uint64, C.c_void_p,
                                      C.c_void_p]),
    "mccfr_tenant_footprint": (C.c_int, [C.POINTER(GameCfg), C.POINTER(C.c_uint64)]),
    "mccfr_tenant_create": (C.c_int, [C.c_void_p, C.POINTER(GameCfg), C.POINTER(RootState), C.POINTER(C.c_uint32)]),
    "mccfr_tenant_destroy": (C.c_int, [C.c_void_p, C.c_uint32]),
    "mccfr_run_segments": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "mccfr_train_tenants": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int64]),
    "mccfr_tenant_counters": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "mccfr_tenant_table_shape": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "mccfr_tenant_download": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                        C.c_int]),
    "mccfr_tenant_upload": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "mccfr_tenant_normalized_strategy": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t, C.c_int]),
    "mccfr_tenant_randomize": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64]),
    "mccfr_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_float)]),
    "mccfr_server_create": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]),
    "mccfr_server_destroy": (None, [C.c_void_p]),
    "mccfr_server_submit": (C.c_int, [C.c_void_p, C.POINTER(GameCfg), C.POINTER(RootState), C.c_int64, C.c_uint32,
                                      C.POINTER(C.c_uint64)]),
    "mccfr_server_step": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "mccfr_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    "mccfr_server_tenant": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32)]),
    "mccfr_server_result": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_size_t, C.c_int]),
    "mccfr_server_release": (C.c_int, [C.c_void_p, C.c_uint64]),
}


def lib() -> C.CDLL:
    """Loads libmccfr.so (raises MccfrError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MccfrError(f"{LIB_PATH} is missing: build it with `make -C {HERE}`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> Sequence[str]:
    return list(_SIGS)


def _check(rc: int, ctx=None) -> None:
    if rc != 0:
        msg = lib().mccfr_last_error(ctx)
        raise MccfrError(f"libmccfr error {rc}: {msg.decode() if msg else ''}", rc)


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------------------
# Game configurations
# ---------------------------------------------------------------------------
def reference_cfg(num_players=2, starting_stack=1000, small_blind=10, big_blind=20, use_cfr_plus=True,
                  use_linear_cfr=True, table_log2=26, seed=42) -> GameCfg:
    """The reference game (GPUCFRTrainer defaults, cfr_gpu.cuh:50-55) with its 8-action menu."""
    cfg = GameCfg()
    lib().mccfr_game_cfg_reference(C.byref(cfg), num_players, starting_stack, small_blind, big_blind,
                                   int(use_cfr_plus), int(use_linear_cfr))
    cfg.table_log2 = table_log2
    cfg.seed = seed
    return cfg


def menu_cfg(num_players, starting_stack, menu, max_raises=0, table_log2=26, seed=0, small_blind=10,
             big_blind=20, use_cfr_plus=True, use_linear_cfr=True, history_bits=None) -> GameCfg:
    """A game with a custom menu: list of (kind, num, den)."""
    cfg = GameCfg()
    cfg.num_players = num_players
    cfg.starting_stack = starting_stack
    cfg.small_blind = small_blind
    cfg.big_blind = big_blind
    cfg.use_cfr_plus = int(use_cfr_plus)
    cfg.use_linear_cfr = int(use_linear_cfr)
    cfg.num_actions = len(menu)
    cfg.max_raises = max_raises
    cfg.history_bits = history_bits if history_bits else max(1, (len(menu) - 1).bit_length())
    cfg.table_log2 = table_log2
    cfg.seed = seed
    for i, (k, n, d) in enumerate(menu):
        cfg.menu[i] = Action(k, n, d, 0)
    return cfg


# Menus of the benchmark configurations.
MENU_FCPA = [(FOLD, 0, 1), (CHECK_CALL, 0, 1), (POT_FRAC, 1, 1), (ALL_IN, 0, 1)]
MENU_6 = [(FOLD, 0, 1), (CHECK_CALL, 0, 1), (POT_FRAC, 1, 2), (POT_FRAC, 1, 1), (POT_FRAC, 2, 1), (ALL_IN, 0, 1)]
MENU_10 = [(FOLD, 0, 1), (CHECK_CALL, 0, 1), (POT_FRAC, 1, 4), (POT_FRAC, 1, 2), (POT_FRAC, 3, 4),
           (POT_FRAC, 1, 1), (POT_FRAC, 3, 2), (POT_FRAC, 2, 1), (POT_FRAC, 3, 1), (ALL_IN, 0, 1)]


def cfg_3p_20bb(table_log2=22, seed=0) -> GameCfg:
    """3 players, 20bb (400 chips at 10/20), {fold, check/call, pot, all-in}, <= 2 raises per street."""
    return menu_cfg(3, 400, MENU_FCPA, max_raises=2, table_log2=table_log2, seed=seed)


def cfg_6p_100bb(table_log2=26, seed=0) -> GameCfg:
    """6 players, 100bb (2,000 chips), {fold, check/call, 1/2, 1, 2 pot, all-in}, <= 3 raises per street."""
    return menu_cfg(6, 2000, MENU_6, max_raises=3, table_log2=table_log2, seed=seed)


def cfg_6p_200bb_10(table_log2=28, seed=0) -> GameCfg:
    """6 players, 200bb (4,000 chips), 10-action menu, <= 4 raises per street."""
    return menu_cfg(6, 4000, MENU_10, max_raises=4, table_log2=table_log2, seed=seed)


# ---------------------------------------------------------------------------
# Hand table
# ---------------------------------------------------------------------------
_table_cache: Optional[np.ndarray] = None


def validate_hand_table(hr: np.ndarray) -> int:
    """Host-only check of a hand table (size, every word in [0, words - 53],
    royal flush = 7462); returns the status code (0 = accepted)."""
    hr = np.ascontiguousarray(hr, np.int32)
    return lib().mccfr_validate_hand_table(_ptr(hr), hr.size)


def validate_hand_table_file(path: str) -> int:
    """Host-only check of a hand-table file (size before any allocation, then
    the contents); returns the status code (0 = accepted)."""
    return lib().mccfr_validate_hand_table_file(path.encode())


def generate_hand_table(threads: int = 0) -> np.ndarray:
    """The 32,487,834-word rank-walk table (generated once per process)."""
    global _table_cache
    if _table_cache is None:
        hr = np.empty(HAND_TABLE_WORDS, dtype=np.int32)
        _check(lib().mccfr_generate_hand_table(_ptr(hr), hr.size, threads))
        _table_cache = hr
    return _table_cache


def preflop_lut() -> np.ndarray:
    out = np.empty(2704, dtype=np.uint8)
    _check(lib().mccfr_preflop_lut(_ptr(out)))
    return out


# ---------------------------------------------------------------------------
# Context
# ---------------------------------------------------------------------------
class Context:
    """Owns one mccfr_ctx (one game, its device tables and stream)."""

    def __init__(self, cfg: GameCfg, max_batch: int):
        self.cfg = cfg
        h = C.c_void_p()
        _check(lib().mccfr_create(C.byref(cfg), max_batch, C.byref(h)))
        self.h = h
        self.max_batch = max_batch
        s = C.c_uint64()
        r = C.c_uint32()
        _check(lib().mccfr_table_shape(h, C.byref(s), C.byref(r)), h)
        self.slots, self.row = int(s.value), int(r.value)

    def close(self):
        if self.h:
            lib().mccfr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, rc):
        _check(rc, self.h)

    # hand table -----------------------------------------------------------
    def use_generated_hand_table(self):
        self.check(lib().mccfr_use_generated_hand_table(self.h))

    def set_hand_table(self, hr: np.ndarray):
        hr = np.ascontiguousarray(hr, dtype=np.int32)
        self.check(lib().mccfr_set_hand_table(self.h, _ptr(hr), hr.size))

    def load_hand_table(self, path: str):
        self.check(lib().mccfr_load_hand_table(self.h, path.encode()))

    # evaluation ------------------------------------------------------------
    def eval7(self, cards: np.ndarray, evaluator: int = 0) -> np.ndarray:
        cards = np.ascontiguousarray(cards, dtype=np.uint8)
        n, stride = cards.shape
        out = np.empty(n, dtype=np.uint16)
        self.check(lib().mccfr_eval7_host(self.h, _ptr(cards), stride, _ptr(out), n, evaluator))
        return out

    # launches --------------------------------------------------------------
    def run_launch(self, update_player: int, t_weight: int, first_tid: int, n: int, trace: bool = False):
        tr = np.zeros(n, dtype=TRACE_DTYPE) if trace else None
        self.check(lib().mccfr_run_launch(self.h, update_player, t_weight, first_tid, n,
                                          _ptr(tr) if trace else None))
        return tr

    def run_launch_noalloc(self, update_player: int, t_weight: int, first_tid: int, n: int):
        """mccfr_run_launch on the library's static scratch (n <= 16384, no trace)."""
        self.check(lib().mccfr_run_launch_noalloc(self.h, update_player, t_weight, first_tid, n))

    def train(self, iters: int, batch: int):
        self.check(lib().mccfr_train(self.h, iters, batch))

    def set_max_batch(self, n: int):
        self.check(lib().mccfr_set_max_batch(self.h, n))
        self.max_batch = n

    def set_launch_shape(self, walk_wg: int, traversals_per_lane: int):
        self.check(lib().mccfr_set_launch_shape(self.h, walk_wg, traversals_per_lane))

    def launch_shape(self):
        w, t = C.c_uint32(), C.c_uint32()
        self.check(lib().mccfr_get_launch_shape(self.h, C.byref(w), C.byref(t)))
        return int(w.value), int(t.value)

    def randomize_tables(self, seed: int):
        self.check(lib().mccfr_randomize_tables(self.h, seed))

    def save_checkpoint(self, path: str):
        self.check(lib().mccfr_save_checkpoint(self.h, path.encode()))

    def load_checkpoint(self, path: str):
        self.check(lib().mccfr_load_checkpoint(self.h, path.encode()))

    def sync(self):
        self.check(lib().mccfr_sync(self.h))

    def download(self, layout: int = 0):
        n = self.slots * self.row
        R = np.empty(n, np.float32)
        S = np.empty(n, np.float32)
        sig = np.empty(n, np.uint16)
        self.check(lib().mccfr_download_tables(self.h, _ptr(R), _ptr(S), _ptr(sig), n, layout))
        return R, S, sig

    def upload(self, R: Optional[np.ndarray], S: Optional[np.ndarray], layout: int = 0):
        n = self.slots * self.row
        Rc = None if R is None else np.ascontiguousarray(R, dtype=np.float32)
        Sc = None if S is None else np.ascontiguousarray(S, dtype=np.float32)
        self.check(lib().mccfr_upload_tables(self.h, None if Rc is None else _ptr(Rc),
                                             None if Sc is None else _ptr(Sc), n, layout))

    def normalized_strategy(self, layout: int = 0) -> np.ndarray:
        n = self.slots * self.row
        out = np.empty(n, np.float32)
        self.check(lib().mccfr_normalized_strategy(self.h, _ptr(out), n, layout))
        return out

    def counters(self):
        t, l, i = C.c_uint64(), C.c_uint64(), C.c_int64()
        self.check(lib().mccfr_counters(self.h, C.byref(t), C.byref(l), C.byref(i)))
        return int(t.value), int(l.value), int(i.value)

    def set_timing(self, on: bool):
        self.check(lib().mccfr_set_timing(self.h, int(on)))

    def kernel_times(self, n: int):
        """(deal_ms, walk_ms, update_ms) arrays of the last n timed launches."""
        a = np.empty(n, np.float32)
        b = np.empty(n, np.float32)
        d = np.empty(n, np.float32)
        self.check(lib().mccfr_kernel_times(self.h, _ptr(a), _ptr(b), _ptr(d), n))
        return a, b, d

    def evaluate_strategy(self, n_players: int, opponent: int, n_hands: int, seed: int = 42, tenant: int = 0,
                          per_hand: bool = False):
        """evaluate_strategy_np (eval.cpp:581-627) on the device: EvalResult (and seat 0's
        net chips per hand with per_hand=True)."""
        res = EvalResult()
        hands = np.empty(n_hands, np.int32) if per_hand else None
        self.check(lib().mccfr_evaluate_strategy(self.h, tenant, n_players, opponent, n_hands, seed,
                                                 _ptr(hands) if per_hand else None, C.byref(res)))
        return (res, hands) if per_hand else res

    def evaluate_rows(self, cfg: GameCfg, keys, probs, n_players: int, opponent: int, n_hands: int,
                      seed: int = 42, per_hand: bool = False):
        """The harness over a strategy given as rows (as load_strategy returns them)."""
        keys = np.ascontiguousarray(keys, np.uint32)
        probs = np.ascontiguousarray(probs, np.float32)
        res = EvalResult()
        hands = np.empty(n_hands, np.int32) if per_hand else None
        self.check(lib().mccfr_evaluate_rows(self.h, C.byref(cfg), _ptr(keys), _ptr(probs), keys.size,
                                             probs.shape[1] if probs.ndim == 2 else 8, n_players, opponent,
                                             n_hands, seed, _ptr(hands) if per_hand else None, C.byref(res)))
        return (res, hands) if per_hand else res

    def save_strategy(self, path: str, tenant: int = 0):
        self.check(lib().mccfr_save_strategy(self.h, tenant, path.encode()))

    def last_launch_stats(self):
        """(records, long runs, pieces) of the last launch."""
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self.check(lib().mccfr_last_launch_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return int(a.value), int(b.value), int(c.value)

    def copy_probe(self, nbytes: int, reps: int) -> float:
        ms = C.c_float()
        self.check(lib().mccfr_copy_probe(self.h, nbytes, reps, C.byref(ms)))
        return float(ms.value)

    # device memory -----------------------------------------------------------
    def gather_probe(self, table_bytes: int, reads: int, reps: int) -> float:
        ms = C.c_float()
        self.check(lib().mccfr_gather_probe(self.h, table_bytes, reads, reps, C.byref(ms)))
        return float(ms.value)

    def dalloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self.check(lib().mccfr_device_alloc(self.h, nbytes, C.byref(p)))
        return int(p.value)

    def dfree(self, p: int):
        self.check(lib().mccfr_device_free(self.h, C.c_void_p(p)))

    def h2d(self, dst: int, src: np.ndarray):
        self.check(lib().mccfr_memcpy(self.h, C.c_void_p(dst), _ptr(src), src.nbytes, 0))

    def d2h(self, dst: np.ndarray, src: int):
        self.check(lib().mccfr_memcpy(self.h, _ptr(dst), C.c_void_p(src), dst.nbytes, 1))

    def random_hands_device(self, seed: int, tag: int, dptr: int, n: int):
        self.check(lib().mccfr_random_hands_device(self.h, seed, tag, C.c_void_p(dptr), n))

    def eval7_device(self, dcards: int, stride: int, dranks: int, n: int, evaluator: int):
        self.check(lib().mccfr_eval7_device(self.h, C.c_void_p(dcards), stride, C.c_void_p(dranks), n, evaluator))

    # tenants and batched launches ---------------------------------------------
    def tenant_create(self, cfg: GameCfg, root: Optional[RootState] = None) -> int:
        t = C.c_uint32()
        self.check(lib().mccfr_tenant_create(self.h, C.byref(cfg), C.byref(root) if root is not None else None,
                                             C.byref(t)))
        return int(t.value)

    def tenant_destroy(self, tenant: int):
        self.check(lib().mccfr_tenant_destroy(self.h, tenant))

    def tenant_shape(self, tenant: int):
        s, r = C.c_uint64(), C.c_uint32()
        self.check(lib().mccfr_tenant_table_shape(self.h, tenant, C.byref(s), C.byref(r)))
        return int(s.value), int(r.value)

    def run_segments(self, segs, trace: bool = False):
        """segs: list of (tenant, n, update_player, t_weight, first_tid)."""
        arr = (Segment * len(segs))()
        for i, (t, n, up, w, tid) in enumerate(segs):
            arr[i] = Segment(t, n, up, 0, w, tid)
        total = sum(s[1] for s in segs)
        tr = np.zeros(total, dtype=TRACE_DTYPE) if trace else None
        self.check(lib().mccfr_run_segments(self.h, C.cast(arr, C.c_void_p), len(segs), _ptr(tr) if trace else None))
        return tr

    def train_tenants(self, tenants, counts, launches: int):
        t = np.ascontiguousarray(tenants, np.uint32)
        c = np.ascontiguousarray(counts, np.uint32)
        self.check(lib().mccfr_train_tenants(self.h, _ptr(t), _ptr(c), t.size, launches))

    def tenant_counters(self, tenant: int):
        a, b, i, p = C.c_uint64(), C.c_uint64(), C.c_int64(), C.c_int32()
        self.check(lib().mccfr_tenant_counters(self.h, tenant, C.byref(a), C.byref(b), C.byref(i), C.byref(p)))
        return int(a.value), int(b.value), int(i.value), int(p.value)

    def tenant_download(self, tenant: int, layout: int = 0):
        slots, row = self.tenant_shape(tenant)
        n = slots * row
        R = np.empty(n, np.float32)
        S = np.empty(n, np.float32)
        sig = np.empty(n, np.uint16)
        self.check(lib().mccfr_tenant_download(self.h, tenant, _ptr(R), _ptr(S), _ptr(sig), n, layout))
        return R, S, sig

    def tenant_upload(self, tenant: int, R, S, layout: int = 0):
        slots, row = self.tenant_shape(tenant)
        n = slots * row
        Rc = None if R is None else np.ascontiguousarray(R, dtype=np.float32)
        Sc = None if S is None else np.ascontiguousarray(S, dtype=np.float32)
        self.check(lib().mccfr_tenant_upload(self.h, tenant, None if Rc is None else _ptr(Rc),
                                             None if Sc is None else _ptr(Sc), n, layout))

    def tenant_normalized_strategy(self, tenant: int, layout: int = 0) -> np.ndarray:
        slots, row = self.tenant_shape(tenant)
        out = np.empty(slots * row, np.float32)
        self.check(lib().mccfr_tenant_normalized_strategy(self.h, tenant, _ptr(out), out.size, layout))
        return out

    def tenant_randomize(self, tenant: int, seed: int):
        self.check(lib().mccfr_tenant_randomize(self.h, tenant, seed))

    def autotune(self, segs, reps: int = 3):
        arr = (Segment * len(segs))()
        for i, (t, n, up, w, tid) in enumerate(segs):
            arr[i] = Segment(t, n, up, 0, w, tid)
        wg, tpl, ms = C.c_uint32(), C.c_uint32(), C.c_float()
        self.check(lib().mccfr_autotune(self.h, C.cast(arr, C.c_void_p), len(segs), reps, C.byref(wg),
                                        C.byref(tpl), C.byref(ms)))
        return int(wg.value), int(tpl.value), float(ms.value)


def load_strategy(path: str):
    """(table_size, num_actions, keys, probs) of a .bin v2 strategy file."""
    ts, na, n = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _check(lib().mccfr_load_strategy(path.encode(), C.byref(ts), C.byref(na), C.byref(n), None, None, 0))
    keys = np.empty(n.value, np.uint32)
    probs = np.empty((n.value, na.value), np.float32)
    _check(lib().mccfr_load_strategy(path.encode(), C.byref(ts), C.byref(na), C.byref(n), _ptr(keys), _ptr(probs),
                                     n.value))
    return int(ts.value), int(na.value), keys, probs


def write_strategy(path: str, table_size: int, keys, probs):
    """Writes rows (keys[n], probs[n, num_actions]) as a .bin v2 strategy file."""
    keys = np.ascontiguousarray(keys, np.uint32)
    probs = np.ascontiguousarray(probs, np.float32)
    if probs.ndim != 2 or probs.shape[0] != keys.shape[0]:
        raise ValueError("probs must be (len(keys), num_actions)")
    _check(lib().mccfr_write_strategy(path.encode(), table_size, probs.shape[1], keys.shape[0], _ptr(keys),
                                      _ptr(probs)))


def row_totals(P: np.ndarray) -> np.ndarray:
    """Float32 row sums added left to right, as the reference's host loop adds them."""
    t = P[:, 0].astype(np.float32)
    for a in range(1, P.shape[1]):
        t = (t + P[:, a]).astype(np.float32)
    return t


def strategy_save(strat: dict, path: str, table_size: int = 1 << 26) -> bool:
    """strategy.cpp:35-52 for a {key: probs} table; rows in ascending key order."""
    keys = np.array(sorted(strat), np.uint32)
    row = len(next(iter(strat.values()))) if strat else 8
    probs = np.array([strat[int(k)] for k in keys], np.float32).reshape(len(keys), row)
    try:
        write_strategy(path, table_size, keys, probs)
    except MccfrError:
        return False
    return True


def strategy_load(path: str) -> dict:
    """strategy.cpp:54-110 without the build-constant check: {key: probs}."""
    _, _, keys, probs = load_strategy(path)
    return {int(k): probs[i] for i, k in enumerate(keys)}


def tenant_footprint(cfg: GameCfg) -> int:
    b = C.c_uint64()
    _check(lib().mccfr_tenant_footprint(C.byref(cfg), C.byref(b)))
    return int(b.value)


WAITING, RUNNING, DONE, RELEASED, FAILED = range(5)


class Server:
    """mccfr_server: continuous batching of subgame re-solves on a Context."""

    def __init__(self, ctx: Context, hbm_budget: int, max_running: int = 256):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(lib().mccfr_server_create(ctx.h, hbm_budget, max_running, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().mccfr_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, cfg: GameCfg, root: Optional[RootState], iterations: int, batch: int) -> int:
        r = C.c_uint64()
        self.ctx.check(lib().mccfr_server_submit(self.h, C.byref(cfg), C.byref(root) if root is not None else None,
                                                 iterations, batch, C.byref(r)))
        return int(r.value)

    def step(self, wait: bool = False):
        a, b = C.c_uint32(), C.c_uint32()
        self.ctx.check(lib().mccfr_server_step(self.h, int(wait), C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def poll(self, req: int):
        st, tr = C.c_int(), C.c_uint64()
        self.ctx.check(lib().mccfr_server_poll(self.h, req, C.byref(st), C.byref(tr)))
        return int(st.value), int(tr.value)

    def tenant(self, req: int) -> int:
        t = C.c_uint32()
        self.ctx.check(lib().mccfr_server_tenant(self.h, req, C.byref(t)))
        return int(t.value)

    def result(self, req: int, layout: int = 0) -> np.ndarray:
        slots, row = self.ctx.tenant_shape(self.tenant(req))
        out = np.empty(slots * row, np.float32)
        self.ctx.check(lib().mccfr_server_result(self.h, req, _ptr(out), out.size, layout))
        return out

    def release(self, req: int):
        self.ctx.check(lib().mccfr_server_release(self.h, req))


class Trainer:
    """GPUCFRTrainer (poker_cuda/src/cfr_gpu.cuh:48-132) over the C ABI:
    the reference's constructor arguments and defaults, load_hand_table,
    train, get_strategy, export/import_tables, num_info_sets_active."""

    def __init__(self, num_players=2, starting_stack=1000, small_blind=10, big_blind=20, use_cfr_plus=True,
                 use_linear_cfr=True, table_log2=26, batch_cap=65536):
        self.cfg = reference_cfg(num_players, starting_stack, small_blind, big_blind, use_cfr_plus, use_linear_cfr,
                                 table_log2=table_log2)
        self.ctx = Context(self.cfg, batch_cap)

    def load_hand_table(self, path: Optional[str] = None) -> bool:
        try:
            if path is None:
                self.ctx.use_generated_hand_table()
            else:
                self.ctx.load_hand_table(path)
        except MccfrError:
            return False
        return True

    def train(self, total_iterations: int, batch_size: int = 65536):
        if batch_size > self.ctx.max_batch:
            self.ctx.set_max_batch(batch_size)
        self.ctx.train(total_iterations, batch_size)

    def get_strategy(self):
        """Normalised rows whose probabilities add to >= 1e-6 (cfr_gpu.cu:1383-1412): {slot: probs}."""
        P = self.ctx.normalized_strategy().reshape(-1, self.ctx.row)
        keep = np.nonzero(row_totals(P) >= np.float32(1e-6))[0]
        return {int(i): P[i] for i in keep}

    def save_strategy(self, path: str) -> bool:
        """get_strategy() written as a .bin v2 file (main.cpp:371-375 with strategy.cpp:35-52)."""
        try:
            self.ctx.save_strategy(path)
        except MccfrError:
            return False
        return True

    def num_info_sets_active(self) -> int:
        _, S, _ = self.ctx.download()
        return int(np.count_nonzero(S.reshape(-1, self.ctx.row).sum(axis=1) > 0))

    def export_tables(self):
        R, S, _ = self.ctx.download(layout=1)
        return R, S

    def import_tables(self, R, S):
        self.ctx.upload(R, S, layout=1)

    def save_checkpoint(self, path: str) -> bool:
        try:
            self.ctx.save_checkpoint(path)
        except MccfrError:
            return False
        return True

    def load_checkpoint(self, path: str) -> bool:
        try:
            self.ctx.load_checkpoint(path)
        except MccfrError:
            return False
        return True
