"""Pins the oracle's tenant roots (postflop re-solve states, mccfr.h
mccfr_root_state) without a GPU.  The root is project-defined (the reference
always starts preflop, cfr_gpu.cu:931-981), so it is pinned by what the rules
fix: deal invariants, the check-down closed form of a launch on all-zero
tables with side pots built from the root's investments, and chip
conservation over update players."""
import numpy as np
import pytest

import mccfr
import oracle
import workloads

P = oracle.p
O = oracle.lib()


def bf_rank(cards7):
    c = np.ascontiguousarray(cards7, np.uint8)
    return O.orc_bf_class(P(c), 7)


def fixed_roots():
    """Hand-picked roots: side pots, folded investors, all-in seats, every street."""
    ref6 = mccfr.menu_cfg(6, 0, mccfr.MENU_6, max_raises=3, table_log2=14, seed=21)
    ref3 = mccfr.reference_cfg(num_players=3, starting_stack=1000, table_log2=14, seed=22)
    ref2 = mccfr.reference_cfg(num_players=2, starting_stack=1000, table_log2=12, seed=23)
    return [
        # flop, 3 of 6 active, folded seats invested, unequal stacks
        (ref6, mccfr.root_state(1, 2, [0, 3, 5], [900, 0, 0, 400, 0, 1500], [100, 20, 10, 100, 0, 100],
                                [12, 25, 38, 0, 0])),
        # turn, 3 active with an all-in short stack: side pot
        (ref3, mccfr.root_state(2, 0, [0, 1, 2], [0, 600, 800], [150, 300, 300], [51, 13, 26, 39, 0], history=0o1234)),
        # river, heads-up
        (ref2, mccfr.root_state(3, 1, [0, 1], [500, 700], [200, 200], [7, 8, 9, 10, 30])),
        # flop, every active seat all-in: straight to showdown, no decisions
        (ref6, mccfr.root_state(1, 4, [1, 2, 4], [0, 0, 0, 0, 0, 0], [0, 50, 300, 20, 700, 0],
                                [1, 2, 3, 0, 0])),
    ]


def test_preflop_root_deal_is_the_reference_deal():
    cfg = mccfr.reference_cfg(num_players=4, starting_stack=1000, seed=5)
    cards = np.zeros(13, np.uint8)
    d = oracle.C.c_int()
    for tid in range(50):
        O.orc_deal(cfg.seed, tid, 0, 4, P(cards), oracle.C.byref(d))
        hole, board = oracle.deal_root(cfg, None, tid)
        assert np.array_equal(hole, cards[:8]) and np.array_equal(board, cards[8:])


@pytest.mark.parametrize("street", [1, 2, 3])
def test_postflop_deal_invariants(street):
    """Fixed board cards stay in place, dealt cards never repeat or hit the
    board, each hole position is uniform over the remaining deck (chi-square)."""
    cfg = mccfr.reference_cfg(num_players=6, starting_stack=1000, seed=31)
    nb = street + 2
    board = [47, 3, 22, 9, 30][:nb] + [0] * (5 - nb)
    root = mccfr.root_state(street, 0, [0, 1, 2, 3, 4, 5], [1000] * 6, [20] * 6, board)
    remaining = [c for c in range(52) if c not in board[:nb]]
    counts = np.zeros(52, np.int64)
    trials = 12000
    for tid in range(trials):
        hole, b = oracle.deal_root(cfg, root, tid)
        assert list(b[:nb]) == board[:nb]
        allc = list(hole) + list(b)
        assert len(set(allc)) == 17
        counts[hole[3]] += 1
    assert counts[board[:nb]].sum() == 0
    exp = trials / len(remaining)
    chi2 = ((counts[remaining] - exp) ** 2 / exp).sum()
    assert chi2 < 110


def checkdown_expected(cfg, root, lut, hr, up, t_weight, first_tid, n):
    """Launch on all-zero tables from a postflop root: every node takes the
    passive fallback (check, no draw) down to showdown, with the depth guard
    of 20 levels counting street advances; side-pot payoff in FP32 from the
    root's investments (folded seats pay in, never win)."""
    N = cfg.num_players
    row = 8 if cfg.num_actions <= 8 else 16
    slots = 1 << cfg.table_log2
    dR = np.zeros(slots * row)
    dS = np.zeros(slots * row)
    utils = np.zeros(n, np.float32)
    folded = {p for p in range(N) if not (root.active_mask >> p) & 1}
    stacks = [int(root.stack[p]) for p in range(N)]
    inv = [int(root.invested[p]) for p in range(N)]
    allin = {p for p in range(N) if p not in folded and stacks[p] == 0}
    pot = sum(inv)
    check = 1  # CHECK (reference menu) or CHECK_CALL (6-action menu)
    for i in range(n):
        hole, board = oracle.deal_root(cfg, root, first_tid + i)
        hole = hole.reshape(N, 2)
        hb = [int(lut[int(h[0]) * 52 + int(h[1])]) for h in hole]
        bkt = [[0] + [O.orc_board_bucket(P(hr), P(np.ascontiguousarray(h)), P(np.ascontiguousarray(board)), s)
                      for s in (1, 2, 3)] for h in hole]

        def order():
            return [(root.dealer + k) % N for k in range(1, N + 1) if (root.dealer + k) % N not in folded | allin]

        street, depth = int(root.street), 0
        hist = int(root.history) & 0x3FFFFFFF
        queue = order()
        nodes, u = [], None
        while True:
            if depth >= 20:
                u = np.float32(0.0)
                break
            if not queue:
                if street < 3:
                    street += 1
                    queue = order()
                    if queue:
                        depth += 1
                        continue
                break  # showdown
            p = queue.pop(0)
            nodes.append((p, street, hist))
            hist = ((hist << cfg.history_bits) | check) & 0x3FFFFFFF
            depth += 1
        if u is None:
            ranks = [bf_rank(np.concatenate([hole[p], board])) for p in range(N)]
            pay = np.float32(-inv[up])
            prev = 0
            for L in sorted({v for v in inv if v > 0}):
                layer = sum(max(0, min(v, L) - prev) for v in inv)
                elig = [q for q in range(N) if q not in folded and inv[q] >= L]
                best = max([ranks[q] for q in elig], default=0)
                if best > 0:
                    n_win = sum(1 for q in elig if ranks[q] == best)
                    if up not in folded and inv[up] >= L and ranks[up] == best:
                        pay = np.float32(pay + np.float32(np.float32(layer) / np.float32(n_win)))
                prev = L
            u = pay
        utils[i] = u
        for (p, st, h) in nodes:
            if p != up:
                continue
            vm = O.orc_valid_mask(oracle.C.byref(cfg), pot, stacks[p], 0, 0, cfg.big_blind, 0)
            legal = [a for a in range(cfg.num_actions) if vm >> a & 1]
            sig = np.float32(1.0) / np.float32(len(legal))
            q = max(sig, np.float32(0.05))
            slot = O.orc_info_hash(p, hb[p], bkt[p][st], st, h, cfg.table_log2)
            for a in legal:
                rd = np.float32(np.float32(u / q) - u) if a == check else np.float32(np.float32(0.0) - u)
                dR[slot * row + a] += float(rd)
                dS[slot * row + a] += float(np.float32(np.float32(t_weight) * sig))
    return dR, dS, utils


@pytest.mark.parametrize("k", range(4))
def test_postflop_first_launch_closed_form(hand_table, golden_lut, k):
    cfg, root = fixed_roots()[k]
    row = 8
    R = np.zeros((1 << cfg.table_log2) * row, np.float32)
    n = 200
    for up in range(cfg.num_players):
        dR, dS, aR, aS, tr = oracle.replay_launch(cfg, hand_table, golden_lut, R, None, up, 2, 500, n, root=root)
        eR, eS, eu = checkdown_expected(cfg, root, golden_lut, hand_table, up, 2, 500, n)
        assert np.array_equal(tr["utility"], eu), up
        assert np.allclose(dR, eR, rtol=1e-12, atol=1e-9)
        assert np.allclose(dS, eS, rtol=1e-12, atol=1e-9)
    if k == 3:
        assert (tr["decisions"] == 0).all() and (tr["terminal"] == 2).all()


def random_tables(slots, row, num_actions, seed):
    rng = np.random.default_rng(seed)
    R = (rng.standard_normal(slots * row) * 50).astype(np.float32)
    R.reshape(slots, row)[:, num_actions:] = 0
    R[rng.random(R.size) < 0.3] = 0
    return R


@pytest.mark.parametrize("k", range(3))
def test_postflop_chips_balance(hand_table, golden_lut, k):
    """From random tables the path does not depend on the update player, and
    the seats' utilities (folded seats: minus their investment) sum to zero
    except at depth-guard terminals -- the roots here give every folded seat
    at most the smallest active investment, so no layer lacks a winner."""
    cfg, root = fixed_roots()[k]
    row = 8
    R = random_tables(1 << cfg.table_log2, row, cfg.num_actions, 3)
    n = 300
    traces = [oracle.replay_launch(cfg, hand_table, golden_lut, R, None, up, 1, 9, n, root=root)[4]
              for up in range(cfg.num_players)]
    for t in traces[1:]:
        assert np.array_equal(t["path_hash"], traces[0]["path_hash"])
    total = np.sum([t["utility"].astype(np.float64) for t in traces], axis=0)
    ok = traces[0]["terminal"] != 3
    assert np.all(np.abs(total[ok]) <= 1e-2)
    assert len(set(traces[0]["decisions"].tolist())) > 3  # the random tables branch
    if k != 1:  # (root 1 keeps an all-in seat in the hand: one fold cannot end it)
        assert (traces[0]["terminal"] == 1).any()


def test_random_roots_are_valid_and_replay(hand_table, golden_lut):
    """The serving workload's random public states pass the library's root
    validation (tenant footprint) and replay in the oracle."""
    rng = np.random.default_rng(4)
    for i in range(20):
        cfg, root = workloads.random_root(rng, table_log2=12, seed=i)
        R = np.zeros((1 << 12) * 8, np.float32)
        tr = oracle.replay_trace(cfg, hand_table, golden_lut, R, None, workloads_first_active(root), 1, 0, 50,
                                 root=root)
        assert (tr["terminal"] >= 1).all()


def workloads_first_active(root):
    m = int(root.active_mask)
    return (m & -m).bit_length() - 1
