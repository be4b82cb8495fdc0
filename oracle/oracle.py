"""ctypes binding of liboracle.so (oracle.cpp): the CPU restatement of the
reference path.  TEST INFRASTRUCTURE ONLY -- imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_lib = None
_P = C.c_void_p

_SIGS = {
    "orc_to_rm": (C.c_int, [C.c_int]),
    "orc_decode_packed": (C.c_uint16, [C.c_uint16]),
    "orc_walk5": (C.c_uint16, [_P, _P]),
    "orc_walk6": (C.c_uint16, [_P, _P]),
    "orc_walk7": (C.c_uint16, [_P, _P]),
    "orc_walk7_batch": (None, [_P, _P, C.c_size_t, _P, C.c_size_t, C.c_int]),
    "orc_validate_table": (C.c_int, [_P, C.c_size_t]),
    "orc_bf_class_count": (C.c_int, [_P]),
    "orc_bf_class": (C.c_int, [_P, C.c_int]),
    "orc_exhaustive7": (C.c_int, [_P, C.c_int, C.c_int, _P, C.POINTER(C.c_uint64)]),
    "orc_exhaustive5_bf": (C.c_int, [_P]),
    "orc_philox4x32_10": (None, [_P, _P, _P]),
    "orc_draw_u32": (C.c_uint32, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]),
    "orc_u01": (C.c_float, [C.c_uint32]),
    "orc_f32_to_bf16": (C.c_uint16, [C.c_float]),
    "orc_bf16_to_f32": (C.c_float, [C.c_uint16]),
    "orc_preflop_lut": (None, [_P]),
    "orc_info_hash": (C.c_uint32, [C.c_uint8, C.c_uint8, C.c_uint8, C.c_uint8, C.c_uint32, C.c_int]),
    "orc_valid_mask": (C.c_uint32, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_chips": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_passive_fallback": (C.c_int, [_P, C.c_uint32]),
    "orc_board_bucket": (C.c_int, [_P, _P, _P, C.c_int]),
    "orc_deal": (None, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int, _P, C.POINTER(C.c_int)]),
    "orc_replay_launch": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, C.c_int64, C.c_uint64, C.c_uint32, C.c_uint32,
                                    _P, _P, _P, _P, _P]),
    "orc_replay_launch_root": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_int, C.c_int64, C.c_uint64, C.c_uint32,
                                         C.c_uint32, _P, _P, _P, _P, _P]),
    "orc_deal_root": (None, [_P, _P, C.c_uint64, C.c_uint32, _P, _P]),
    "orc_regret_matching": (None, [_P, _P, C.c_uint64, C.c_int]),
    "orc_eval_hands": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                                 C.c_int64, C.c_int, C.c_int, _P, _P]),
    "orc_eval_scripted": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, _P]),
    "orc_eval_scripted_action": (C.c_int, [_P, C.c_int, C.c_uint32, C.c_int, C.c_uint64, C.c_int64, _P]),
    "orc_eval_buckets": (C.c_int, [_P, _P, C.c_int, _P, _P, _P]),
    "orc_replay_sparse": (C.c_int64, [_P, _P, _P, _P, _P, _P, C.c_int, C.c_int64, C.c_uint64, C.c_uint32,
                                      C.c_uint32, C.c_int, _P, _P]),
    "orc_sparse_free": (None, [_P]),
    "orc_walk_cards": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_int, _P, _P, C.c_int, C.c_int64, C.c_uint64, C.c_uint32,
                                 _P, _P, C.c_int, C.POINTER(C.c_int), _P, _P, _P, _P, C.c_int, C.POINTER(C.c_int)]),
    "orc_cpu_baseline": (C.c_double, [_P, _P, _P, _P, _P, _P, C.c_int, C.c_uint32, C.c_int, C.c_int64,
                                      C.c_uint64, C.c_int]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C {HERE}`")
        L = C.CDLL(LIB_PATH)
        for k, (r, a) in _SIGS.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def ref_lib():
    """The reference's own CPU sources compiled under oracle/_ref (None when not built)."""
    if not os.path.exists(REF_PATH):
        return None
    L = C.CDLL(REF_PATH)
    L.ref_hand_eval_init.argtypes = [C.c_char_p]
    L.ref_evaluate_7cards.argtypes = [_P]
    L.ref_evaluate_best.argtypes = [_P, _P, C.c_int]
    L.ref_preflop_bucket.argtypes = [C.c_int, C.c_int]
    L.ref_valid_actions_mask.argtypes = [C.c_int] * 6
    L.ref_action_to_chips.argtypes = [C.c_int] * 7
    if hasattr(L, "ref_precompute_buckets"):
        L.ref_precompute_buckets.argtypes = [_P, _P, C.c_int, _P, _P]
    return L


TRACE_DTYPE = np.dtype([("utility", "<f4"), ("path_hash", "<u4"), ("decisions", "<u2"), ("updates", "u1"),
                        ("terminal", "u1"), ("update_legal", "<u2"), ("reserved", "<u2")])


def _root_ptr(root):
    return None if root is None else C.byref(root)


def replay_trace(cfg, hr, lut, R, sigma, up, t_weight, first_tid, n, tag=0, root=None):
    """Per-traversal records only (no delta accumulation): for full-size tables."""
    tr = np.zeros(n, dtype=TRACE_DTYPE)
    R = np.ascontiguousarray(R, np.float32)
    sig_p = None if sigma is None else p(np.ascontiguousarray(sigma, np.uint16))
    rc = lib().orc_replay_launch_root(C.byref(cfg), _root_ptr(root), p(hr), p(lut), p(R), sig_p, up, t_weight,
                                      first_tid, n, tag, None, None, None, None, p(tr))
    if rc != 0:
        raise RuntimeError("orc_replay_launch_root rejected its arguments")
    return tr


def replay_launch(cfg, hr, lut, R, sigma, up, t_weight, first_tid, n, tag=0, trace=True, root=None):
    """FP64 replay of one launch from pre-launch tables (slot-major rows).
    Returns (dR, dS, aR, aS, trace)."""
    row = 8 if cfg.num_actions <= 8 else 16
    size = (1 << cfg.table_log2) * row
    dR = np.zeros(size); dS = np.zeros(size); aR = np.zeros(size); aS = np.zeros(size)
    tr = np.zeros(n, dtype=TRACE_DTYPE) if trace else None
    R = np.ascontiguousarray(R, np.float32)
    sig_p = None if sigma is None else p(np.ascontiguousarray(sigma, np.uint16))
    rc = lib().orc_replay_launch_root(C.byref(cfg), _root_ptr(root), p(hr), p(lut), p(R), sig_p, up, t_weight,
                                      first_tid, n, tag, p(dR), p(dS), p(aR), p(aS), p(tr) if trace else None)
    if rc != 0:
        raise RuntimeError("orc_replay_launch_root rejected its arguments")
    return dR, dS, aR, aS, tr


DECISION_DTYPE = np.dtype([("player", "u1"), ("street", "u1"), ("action", "u1"), ("is_update", "u1"),
                           ("legal", "<u4"), ("history", "<u4"), ("slot", "<u4"), ("chips", "<i4"), ("pot", "<i4"),
                           ("current_bet", "<i4"), ("lfr", "<i4"), ("raises", "<i4"), ("n_queue", "<i4"),
                           ("queue", "<u4")])


def walk_cards(cfg, hr, lut, hole, board, dealer, R, sigma, up, t_weight=1, tid=0, tag=0, root=None, cap=64):
    """One traversal with given cards and dealer: (trace record, decision log,
    contributions as (slot, action, regret delta, strategy-sum delta) arrays)."""
    hole = np.ascontiguousarray(hole, np.uint8)
    board = np.ascontiguousarray(board, np.uint8)
    R = np.ascontiguousarray(R, np.float32)
    sig_p = None if sigma is None else p(np.ascontiguousarray(sigma, np.uint16))
    tr = np.zeros(1, dtype=TRACE_DTYPE)
    log = np.zeros(cap, dtype=DECISION_DTYPE)
    cs = np.zeros(cap * 16, np.uint32)
    ca = np.zeros(cap * 16, np.uint32)
    crd = np.zeros(cap * 16, np.float32)
    csd = np.zeros(cap * 16, np.float32)
    n_log, n_c = C.c_int(), C.c_int()
    rc = lib().orc_walk_cards(C.byref(cfg), _root_ptr(root), p(hr), p(lut), p(hole), p(board), dealer, p(R), sig_p,
                              up, t_weight, tid, tag, p(tr), p(log), cap, C.byref(n_log), p(cs), p(ca), p(crd),
                              p(csd), cap * 16, C.byref(n_c))
    if rc != 0:
        raise RuntimeError("orc_walk_cards rejected its arguments")
    k = n_c.value
    return tr[0], log[: n_log.value], (cs[:k], ca[:k], crd[:k], csd[:k])


class _Sparse(C.Structure):
    _fields_ = [("n", C.c_int64), ("idx", C.POINTER(C.c_uint64)), ("dR", C.POINTER(C.c_double)),
                ("dS", C.POINTER(C.c_double)), ("aR", C.POINTER(C.c_double)), ("aS", C.POINTER(C.c_double))]


def replay_sparse(cfg, hr, lut, R, sigma, up, t_weight, first_tid, n, tag=0, root=None, threads=16, trace=True):
    """Multithreaded FP64 replay of one launch for full-size tables: (idx, dR,
    dS, aR, aS) over the touched entries (idx = slot * row + action), and the
    per-traversal records."""
    R = np.ascontiguousarray(R, np.float32)
    sigma = np.ascontiguousarray(sigma, np.uint16)
    tr = np.zeros(n, dtype=TRACE_DTYPE) if trace else None
    sp = _Sparse()
    k = lib().orc_replay_sparse(C.byref(cfg), _root_ptr(root), p(hr), p(lut), p(R), p(sigma), up, t_weight,
                                first_tid, n, tag, threads, p(tr) if trace else None, C.byref(sp))
    if k < 0:
        raise RuntimeError("orc_replay_sparse failed")
    try:
        out = tuple(np.ctypeslib.as_array(getattr(sp, f), shape=(k,)).copy() if k else np.zeros(0)
                    for f in ("idx", "dR", "dS", "aR", "aS"))
    finally:
        lib().orc_sparse_free(C.byref(sp))
    return out + (tr,)


def deal_root(cfg, root, tid, tag=0):
    """(hole cards [2N], board [5]) of traversal tid from a root (None = preflop)."""
    hole = np.zeros(2 * cfg.num_players, np.uint8)
    board = np.zeros(5, np.uint8)
    lib().orc_deal_root(C.byref(cfg), _root_ptr(root), tid, tag, p(hole), p(board))
    return hole, board


def regret_matching(R, row):
    R = np.ascontiguousarray(R, np.float32)
    out = np.empty(R.size, np.uint16)
    lib().orc_regret_matching(p(R), p(out), R.size // row, row)
    return out


# ---- strategy-quality harness (eval.cpp:45-627) ----
CALLING_STATION, NIT, MANIAC, BALANCED, RANDOM = range(5)


def eval_hands(cfg, hr, lut, P, n_players, opponent, n_hands, seed=42, first_hand=0, stream=0, threads=1,
               stack=None, seats=False):
    """Net chips of seat 0 per hand (and of every seat with seats=True)."""
    row = 8 if cfg.num_actions <= 8 else 16
    P = np.ascontiguousarray(P, np.float32)
    bot = np.zeros(n_hands, np.int32)
    seat = np.zeros((n_hands, n_players), np.int32) if seats else None
    st = cfg.starting_stack if stack is None else stack
    rc = lib().orc_eval_hands(C.byref(cfg), p(hr), p(lut), p(P), row, n_players, opponent, st, seed, first_hand,
                              n_hands, stream, threads, p(bot), p(seat) if seats else None)
    if rc != 0:
        raise RuntimeError("orc_eval_hands rejected its arguments")
    return (bot, seat) if seats else bot


def eval_scripted(cfg, hr, lut, n_players, dealer, cards, actions, stack=None):
    """(decisions used, net chips per seat) of one hand with given cards and decisions."""
    cards = np.ascontiguousarray(cards, np.uint8)
    acts = np.ascontiguousarray(actions, np.int32)
    net = np.zeros(n_players, np.int32)
    st = cfg.starting_stack if stack is None else stack
    used = lib().orc_eval_scripted(C.byref(cfg), p(hr), p(lut), n_players, st, dealer, p(cards), p(acts),
                                   len(acts), p(net))
    return used, net


def eval_scripted_counts(cfg, opponent, vm, to_call, draws, seed=1):
    counts = np.zeros(16, np.int64)
    rc = lib().orc_eval_scripted_action(C.byref(cfg), opponent, vm, to_call, seed, draws, p(counts))
    if rc != 0:
        raise RuntimeError("orc_eval_scripted_action rejected its arguments")
    return counts


def eval_buckets(hr, lut, n_players, cards):
    cards = np.ascontiguousarray(cards, np.uint8)
    hb = np.zeros(n_players, np.int32)
    bk = np.zeros((n_players, 4), np.int32)
    lib().orc_eval_buckets(p(hr), p(lut), n_players, p(cards), p(hb), p(bk))
    return hb, bk
