"""Pins the CPU oracle (oracle/oracle.cpp) against fixtures, known answers and
mathematics, without a GPU.  Each test names what it pins."""
import ctypes as C
import os

import numpy as np
import pytest

import helpers
import mccfr
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = oracle.p
O = oracle.lib()


def rm(c):  # rank-major index, hand_eval.cpp:23-27
    return (c % 13) * 4 + c // 13 + 1


def np_walk(hr, cards):
    """Vectorised rank walk in the given card order; returns the word after each step."""
    p = hr[53 + rm(cards[:, 0].astype(np.int64))].astype(np.int64)
    out = [p]
    for k in range(1, cards.shape[1]):
        p = hr[p + rm(cards[:, k].astype(np.int64))].astype(np.int64)
        out.append(p)
    return out


def decode(packed):
    packed = np.asarray(packed, np.int64) & 0xFFFF
    cat, ordv = packed >> 12, packed & 0xFFF
    off = np.array([0, 0, 1277, 4137, 4995, 5853, 5863, 7140, 7296, 7452, 0, 0, 0, 0, 0, 0])
    return np.where((ordv == 0) | (cat < 1) | (cat > 9), 0, off[cat] + ordv)


# ---------------------------------------------------------------------------
# Hand table and evaluators
# ---------------------------------------------------------------------------
def test_table_size_and_royal_kat(hand_table):
    """hand_eval.cpp:14 size and :70-84 royal flush = 7462."""
    assert hand_table.size == 32487834
    assert O.orc_validate_table(P(hand_table), hand_table.size) == 1
    royal = np.array([51, 50, 49, 48, 47, 0, 14], np.uint8)
    assert O.orc_walk7(P(hand_table), P(royal)) == 7462


def test_bruteforce_class_counts():
    """7462 distinct 5-card classes split 1277/2860/858/858/10/1277/156/156/10 (textbook)."""
    per = np.zeros(9, np.int32)
    assert O.orc_bf_class_count(P(per)) == 7462
    assert per.tolist() == [1277, 2860, 858, 858, 10, 1277, 156, 156, 10]


def test_exhaustive_five_card_counts():
    """All 2,598,960 five-card hands by category (textbook frequencies)."""
    c = np.zeros(10, np.uint64)
    O.orc_exhaustive5_bf(P(c))
    assert c[1:].tolist() == [1302540, 1098240, 123552, 54912, 10200, 5108, 3744, 624, 40]


def test_exhaustive_seven_card_table_vs_bruteforce(hand_table):
    """All 133,784,560 seven-card hands: walk categories match the textbook
    counts and every walk rank equals the brute-force best-of-21 class."""
    c = np.zeros(10, np.uint64)
    mm = C.c_uint64()
    O.orc_exhaustive7(P(hand_table), 0, 1, P(c), C.byref(mm))
    assert c[1:].tolist() == [23294460, 58627800, 31433400, 6461620, 6180020, 4047644, 3473184, 224848, 41584]
    assert mm.value == 0


def test_walks_match_reference_code(hand_table):
    """The reference's evaluate_best / evaluate_7cards (hand_eval.cpp, run on this
    table by tests/golden/make_golden.py) equal the oracle's walks, including
    the 5- and 6-card walks that decode a state pointer (cfr_gpu.cu:114-150)."""
    g = np.load(os.path.join(GOLDEN, "hand_walks.npz"))
    cards, ranks = g["cards"], g["ranks"]
    for c, r in zip(cards, ranks):
        c = np.ascontiguousarray(c)
        assert O.orc_walk5(P(hand_table), P(c)) == r[0]
        assert O.orc_walk6(P(hand_table), P(c)) == r[1]
        assert O.orc_walk7(P(hand_table), P(c)) == r[2] == r[3]
        assert O.orc_bf_class(P(c), 7) == r[3]


def test_walk_state_is_order_free(hand_table):
    """The walk state after k cards depends only on the card set: walking the
    3 flop cards first and the hole cards after gives the same 5-, 6- and
    7-card words as the reference's hole-first order.  The kernel relies on
    this to share the board prefix across players."""
    rng = np.random.default_rng(7)
    cards = np.stack([rng.permutation(52)[:7] for _ in range(200000)]).astype(np.int64)
    ref_order = np_walk(hand_table, cards)
    board_first = np_walk(hand_table, cards[:, [2, 3, 4, 0, 1, 5, 6]])
    for k in (4, 5, 6):
        assert np.array_equal(ref_order[k], board_first[k])
    g = np.load(os.path.join(GOLDEN, "hand_walks.npz"))
    ws = np_walk(hand_table, g["cards"][:, [2, 3, 4, 0, 1, 5, 6]].astype(np.int64))
    assert np.array_equal(decode(ws[4]), g["ranks"][:, 0])
    assert np.array_equal(decode(ws[5]), g["ranks"][:, 1])
    assert np.array_equal(decode(ws[6]), g["ranks"][:, 3])


# ---------------------------------------------------------------------------
# Random streams, rounding, hashing
# ---------------------------------------------------------------------------
def test_philox_matches_rocrand_and_random123():
    """Philox4x32-10 against rocRAND's implementation (256 vectors, incl. the
    Random123 known answers for all-zero and all-ones inputs)."""
    v = np.fromfile(os.path.join(GOLDEN, "philox_vectors.bin"), np.uint32).reshape(-1, 10)
    assert v[0, 6:].tolist() == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert v[1, 6:].tolist() == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    out = np.zeros(4, np.uint32)
    for rec in v:
        ctr = np.ascontiguousarray(rec[0:4])
        key = np.ascontiguousarray(rec[4:6])
        O.orc_philox4x32_10(P(ctr), P(key), P(out))
        assert out.tolist() == rec[6:].tolist()


def test_draw_layout():
    """Draw k of traversal t = word k&3 of Philox({k>>2, t_lo, t_hi, tag}, {seed_lo, seed_hi})."""
    out = np.zeros(4, np.uint32)
    seed, tid, tag = 0x1234_5678_9ABC_DEF0, 0x0000_0003_0000_0010, 7
    for k in range(12):
        ctr = np.array([k >> 2, tid & 0xFFFFFFFF, tid >> 32, tag], np.uint32)
        key = np.array([seed & 0xFFFFFFFF, seed >> 32], np.uint32)
        O.orc_philox4x32_10(P(ctr), P(key), P(out))
        assert O.orc_draw_u32(seed, tid, tag, k) == out[k & 3]


def test_uniform_range():
    assert O.orc_u01(0) == np.float32(2.0 ** -24)
    assert O.orc_u01(0xFFFFFFFF) == 1.0
    assert O.orc_u01(0x80000000) == np.float32(0.5 + 2.0 ** -24)


def test_bf16_round_to_nearest_even():
    """float32 -> bfloat16 RNE checked against an integer-arithmetic definition."""
    rng = np.random.default_rng(3)
    bits = np.concatenate([rng.integers(0, 2 ** 31, 20000, dtype=np.uint64).astype(np.uint32),
                           np.array([0x3F808000, 0x3F818000, 0x3F807FFF, 0x3F800000, 0x00000000], np.uint32)])
    bits = bits[(bits & 0x7F800000) != 0x7F800000]
    f = bits.view(np.float32)
    hi, lo = bits >> 16, bits & 0xFFFF
    want = np.where((lo > 0x8000) | ((lo == 0x8000) & (hi & 1 == 1)), hi + 1, hi).astype(np.uint16)
    got = np.array([O.orc_f32_to_bf16(C.c_float(x)) for x in f], np.uint16)
    assert np.array_equal(got, want)


def test_info_hash_is_fnv1a():
    """FNV-1a-32 over (player, hole_b, board_b, street, 4 history bytes) (cfr_gpu.cu:165-179)."""
    rng = np.random.default_rng(5)
    for _ in range(2000):
        pl, hb, bb, st = (int(x) for x in rng.integers(0, 256, 4))
        ab = int(rng.integers(0, 2 ** 30))
        log2 = int(rng.integers(4, 33))
        h = helpers.fnv1a(bytes([pl, hb, bb, st]) + ab.to_bytes(4, "little"))
        assert O.orc_info_hash(pl, hb, bb, st, ab, log2) == h & ((1 << log2) - 1)


# ---------------------------------------------------------------------------
# Abstraction against the reference's own code (golden fixtures)
# ---------------------------------------------------------------------------
def test_preflop_lut_matches_reference(golden_lut):
    lut = np.zeros(2704, np.uint8)
    O.orc_preflop_lut(P(lut))
    assert np.array_equal(lut, golden_lut)
    assert np.array_equal(mccfr.preflop_lut(), golden_lut)  # the product's own table builder


def test_rules_match_reference():
    """valid_actions_mask / action_to_chips of the 8-action menu over a grid
    (abstraction.cpp:144-234 == cfr_gpu.cu:186-244)."""
    g = np.load(os.path.join(GOLDEN, "rules_grid.npz"))
    for (pot, stack, tc, cb, lfr, bb), m, ch in zip(g["states"], g["masks"], g["chips"]):
        cfg = mccfr.reference_cfg(big_blind=int(bb))
        assert O.orc_valid_mask(C.byref(cfg), int(pot), int(stack), int(tc), int(cb), int(lfr), 0) == m
        for a in range(8):
            assert O.orc_chips(C.byref(cfg), a, int(pot), int(stack), int(tc), int(cb), int(lfr)) == ch[a]


def test_raise_cap_removes_raises():
    cfg = mccfr.cfg_6p_100bb()
    free = O.orc_valid_mask(C.byref(cfg), 100, 1000, 20, 40, 20, 2)
    capped = O.orc_valid_mask(C.byref(cfg), 100, 1000, 20, 40, 20, 3)
    assert free == 0b111111  # fold, call, 1/2, 1, 2 pot, all-in
    assert capped == 0b000011  # fold, call
    assert O.orc_valid_mask(C.byref(cfg), 100, 10, 20, 40, 20, 3) == 0b100001  # short: fold / all-in


def test_deal_invariants():
    """2N+5 distinct cards per deal; each position uniform over 52 (chi-square)."""
    cards = np.zeros(17, np.uint8)
    d = C.c_int()
    counts = np.zeros((17, 52), np.int64)
    dealers = np.zeros(6, np.int64)
    for t in range(20000):
        O.orc_deal(11, t, 0, 6, P(cards), C.byref(d))
        assert len(set(cards.tolist())) == 17
        counts[np.arange(17), cards] += 1
        dealers[d.value] += 1
    exp = 20000 / 52
    chi2 = ((counts - exp) ** 2 / exp).sum(axis=1)
    assert (chi2 < 110).all()  # df = 51, p < 1e-6 beyond 110
    assert ((dealers - 20000 / 6) ** 2 / (20000 / 6)).sum() < 30


# ---------------------------------------------------------------------------
# Traversal
# ---------------------------------------------------------------------------
def bf_rank(cards7):
    c = np.ascontiguousarray(cards7, np.uint8)
    return O.orc_bf_class(P(c), 7)


def first_launch_expected(cfg, lut, hr, up, t_weight, first_tid, n):
    """Closed form of a launch on all-zero tables (cfr_gpu.cu:1117-1119, :520-521):
    every BF16 row is 0, so every node takes the passive fallback without a
    draw.  With stacks above the big blind every seat calls, the big blind
    checks and all check to showdown -- unless the depth guard of 20 levels
    (decisions + street advances, cfr_gpu.cu:642, :661) ends the walk first.
    The update node's sigma is uniform over the legal actions, q = max(1/k, 0.05),
    the sampled action is check or call."""
    N, bb = cfg.num_players, cfg.big_blind
    row = 8 if cfg.num_actions <= 8 else 16
    slots = 1 << cfg.table_log2
    dR = np.zeros(slots * row)
    dS = np.zeros(slots * row)
    utils = np.zeros(n, np.float32)
    cards = np.zeros(2 * N + 5, np.uint8)
    d = C.c_int()
    for i in range(n):
        O.orc_deal(cfg.seed, first_tid + i, 0, N, P(cards), C.byref(d))
        dealer = d.value
        hole = cards[: 2 * N].reshape(N, 2)
        board = cards[2 * N:]
        hb = [int(lut[int(h[0]) * 52 + int(h[1])]) for h in hole]
        bkt = [[0] + [O.orc_board_bucket(P(hr), P(np.ascontiguousarray(h)), P(np.ascontiguousarray(board)), s)
                      for s in (1, 2, 3)] for h in hole]
        sb_pos = dealer if N == 2 else (dealer + 1) % N
        bb_pos = (dealer + 1) % N if N == 2 else (dealer + 2) % N
        stacks = [cfg.starting_stack] * N
        bets = [0] * N
        for pos, amt in ((sb_pos, cfg.small_blind), (bb_pos, bb)):
            stacks[pos] -= amt
            bets[pos] = amt
        pot, cur = cfg.small_blind + bb, bb
        nodes = []  # (player, street, history, pot, stack, to_call, current_bet)
        hist, depth, u = 0, 0, None
        for street in range(4):
            if street == 0:
                order = [(bb_pos + 1 + k) % N for k in range(N)]
            else:
                if depth >= 20:
                    u = 0.0
                    break
                depth += 1  # a street advance is one recursion level
                order = [(dealer + k) % N for k in range(1, N + 1)]
                bets = [0] * N
                cur = 0
            for p in order:
                if depth >= 20:
                    u = 0.0
                    break
                to_call = cur - bets[p]
                nodes.append((p, street, hist, pot, stacks[p], to_call, cur))
                a = 2 if to_call > 0 else 1  # CALL or CHECK of the 8-action menu
                stacks[p] -= to_call
                bets[p] += to_call
                pot += to_call
                hist = ((hist << 3) | a) & 0x3FFFFFFF
                depth += 1
            if u is not None:
                break
        if u is None:
            if depth >= 20:
                u = 0.0
            else:
                ranks = [bf_rank(np.concatenate([hole[p], board])) for p in range(N)]
                best = max(ranks)
                u = np.float32(-bb)
                if ranks[up] == best:
                    u = np.float32(u + np.float32(np.float32(N * bb) / np.float32(ranks.count(best))))
        u = np.float32(u)
        utils[i] = u
        for (p, street, h, pot_n, stack_n, to_call, cur_n) in nodes:
            if p != up:
                continue
            vm = O.orc_valid_mask(C.byref(cfg), pot_n, stack_n, to_call, cur_n, bb, 0)
            legal = [a for a in range(cfg.num_actions) if vm >> a & 1]
            sig = np.float32(1.0) / np.float32(len(legal))
            q = max(sig, np.float32(0.05))
            sampled = 2 if to_call > 0 else 1
            slot = O.orc_info_hash(p, hb[p], bkt[p][street], street, h, cfg.table_log2)
            for a in legal:
                rd = np.float32(np.float32(u / q) - u) if a == sampled else np.float32(np.float32(0.0) - u)
                dR[slot * row + a] += float(rd)
                dS[slot * row + a] += float(np.float32(np.float32(t_weight) * sig))
    return dR, dS, utils


@pytest.mark.parametrize("N,stack", [(2, 1000), (3, 400), (6, 2000)])
def test_first_launch_closed_form(hand_table, golden_lut, N, stack):
    cfg = mccfr.reference_cfg(num_players=N, starting_stack=stack, table_log2=14, seed=9)
    slots, row = 1 << 14, 8
    R = np.zeros(slots * row, np.float32)
    n = 300
    for up in range(N):
        dR, dS, aR, aS, tr = oracle.replay_launch(cfg, hand_table, golden_lut, R, None, up, 3, 1000, n)
        eR, eS, eu = first_launch_expected(cfg, golden_lut, hand_table, up, 3, 1000, n)
        assert np.array_equal(tr["utility"], eu)
        assert np.allclose(dR, eR, rtol=1e-12, atol=1e-9)
        assert np.allclose(dS, eS, rtol=1e-12, atol=1e-9)
        if N == 6:
            assert (tr["terminal"] == 3).all()  # 6-handed check-down reaches the depth guard on the turn


def random_tables(slots, row, num_actions, seed):
    rng = np.random.default_rng(seed)
    R = (rng.standard_normal(slots * row) * 50).astype(np.float32)
    R.reshape(slots, row)[:, num_actions:] = 0
    R[rng.random(R.size) < 0.3] = 0
    return R


@pytest.mark.parametrize("cfg_fn", [
    lambda: mccfr.reference_cfg(num_players=3, starting_stack=400, table_log2=12, seed=4),
    lambda: mccfr.cfg_6p_100bb(table_log2=12, seed=4),
    lambda: mccfr.cfg_3p_20bb(table_log2=12, seed=4),
])
def test_paths_do_not_depend_on_update_player_and_chips_balance(hand_table, golden_lut, cfg_fn):
    """Sampling uses the same snapshot and draws whoever the update player is,
    so the path is the same for every update player; the update players'
    utilities then sum to zero over the seats (chip conservation) except for
    depth-guard terminals (all zero) and FP32 split rounding."""
    cfg = cfg_fn()
    row = 8 if cfg.num_actions <= 8 else 16
    slots = 1 << cfg.table_log2
    R = random_tables(slots, row, cfg.num_actions, 1)
    n = 400
    traces = [oracle.replay_launch(cfg, hand_table, golden_lut, R, None, up, 1, 77, n)[4] for up in
              range(cfg.num_players)]
    for t in traces[1:]:
        assert np.array_equal(t["path_hash"], traces[0]["path_hash"])
        assert np.array_equal(t["decisions"], traces[0]["decisions"])
    total = np.sum([t["utility"].astype(np.float64) for t in traces], axis=0)
    assert np.all(np.abs(total) <= 1e-3 * cfg.starting_stack)
    assert len(set(traces[0]["decisions"].tolist())) > 3  # the random tables do branch


def test_update_node_invariant(hand_table, golden_lut):
    """sum_a sigma(a) * rdelta(a) = 0 at an update node whose q is not floored:
    sigma(a*)(u/q - u) - u(1 - sigma(a*)) = 0 (single-traversal launches, so
    each touched row holds one node's contributions)."""
    cfg = mccfr.reference_cfg(num_players=2, starting_stack=1000, table_log2=12, seed=2)
    slots, row = 1 << 12, 8
    R = np.abs(random_tables(slots, row, 8, 2))
    checked = 0
    for tid in range(300):
        dR, dS, aR, aS, tr = oracle.replay_launch(cfg, hand_table, golden_lut, R, None, 0, 1, tid, 1)
        if tr["updates"][0] != 1:
            continue
        s = np.nonzero(aS.reshape(slots, row).sum(axis=1))[0]
        assert len(s) == 1
        s = s[0]
        sig = dS.reshape(slots, row)[s]  # t = 1: strategy-sum deltas are sigma itself
        rd = dR.reshape(slots, row)[s]
        legal = aS.reshape(slots, row)[s] > 0
        if (sig[legal] < 0.05).any():
            continue
        assert abs(float(np.dot(sig, rd))) <= 1e-5 * float(np.abs(rd).sum() + 1)
        checked += 1
    assert checked > 20


def test_cpu_baseline_matches_replay(hand_table, golden_lut):
    """The timed FP32 baseline (SoA tables, full sweeps) lands within the
    tolerance of pre-launch tables + the FP64 replay's contributions."""
    cfg = mccfr.reference_cfg(num_players=3, starting_stack=400, table_log2=12, seed=5)
    slots, row = 1 << 12, 8
    R0 = np.abs(random_tables(slots, row, 8, 3))
    S0 = np.zeros_like(R0)
    Rs = np.ascontiguousarray(R0.reshape(slots, row).T.reshape(-1))
    Ss = np.zeros_like(Rs)
    sig = np.zeros(slots * row, np.uint16)
    n = 2000
    secs = O.orc_cpu_baseline(C.byref(cfg), P(hand_table), P(golden_lut), P(Rs), P(Ss), P(sig), 1, n, 1, 4, 500, 4)
    assert secs >= 0
    dR, dS, aR, aS, _ = oracle.replay_launch(cfg, hand_table, golden_lut, R0, None, 1, 4, 500, n)
    eR, eS = helpers.expected_post(R0, S0, dR, dS)
    gotR = Rs.reshape(row, slots).T.reshape(-1)
    gotS = Ss.reshape(row, slots).T.reshape(-1)
    # FP32 accumulation in the baseline: a looser bound for it (it is not the product)
    assert np.all(np.abs(gotR - eR) <= 1e-4 * np.maximum(np.abs(eR), aR) + 1e-6)
    assert np.all(np.abs(gotS - eS) <= 1e-4 * np.maximum(np.abs(eS), aS) + 1e-6)
