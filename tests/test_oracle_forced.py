"""Pins the oracle's betting walk (oracle.cpp Walker: rules, apply, queue
rebuild, raise cap, history, side pots, depth guard) against hand-worked
sequences with forced play, and the all-in two-player games against the
brute-force evaluator.  CPU only.

Forced play: a one-hot BF16 strategy row at a node's info-set slot makes the
node take that action whatever the draw (cfr_gpu.cu:510-537: the running sum
crosses r in (0, total] at the one non-zero legal entry); an all-zero row, or
a one-hot row on an illegal action, takes the passive fallback
(cfr_gpu.cu:498-508).  The expected decision logs, utilities and slots are the
literals of tests/forced_play.py."""
import numpy as np
import pytest

import forced_play as fp
import mccfr
import oracle

LOG2 = 16
BF16_ONE = 0x3F80


def setup(sc, hand_table, lut, hole=None, board=None):
    cfg = sc.make_cfg(LOG2)
    row = 8 if cfg.num_actions <= 8 else 16
    hole = np.array(fp.cards(*sc.hole) if hole is None else hole, np.uint8)
    board = np.array(fp.cards(*sc.board) if board is None else board, np.uint8)
    root = fp.root_of(sc, cfg, board)
    slots = fp.node_slots(sc, hole, board, lut, hand_table, LOG2, oracle)
    need = fp.forced_rows(sc, slots, row)
    assert need is not None, "two scripted nodes share a slot"
    sig = np.zeros((1 << LOG2) * row, np.uint16)
    for s, a in need.items():
        if a is not None:
            sig[s * row + a] = BF16_ONE
    return cfg, row, hole, board, root, slots, sig


def legal_of(mask):
    return tuple(a for a in range(16) if mask >> a & 1)


def queue_of(d):
    return tuple((int(d["queue"]) >> (4 * i)) & 15 for i in range(int(d["n_queue"])))


@pytest.mark.parametrize("sc", fp.SCENARIOS, ids=lambda s: s.name)
def test_forced_walk_matches_hand_worked_sequence(sc, hand_table, golden_lut):
    cfg, row, hole, board, root, slots, sig = setup(sc, hand_table, golden_lut)
    R = np.zeros((1 << LOG2) * row, np.float32)  # uniform sigma at update nodes
    ranks = fp.ranks_of(hole, board, oracle)
    u = fp.utilities(sc, ranks)
    assert u is not None, "the scenario's cards fit none of its cases"
    taken = [nd.taken for nd in sc.nodes]
    tw = 3
    for up in range(cfg.num_players):
        tr, log, (cs, ca, crd, csd) = oracle.walk_cards(cfg, hand_table, golden_lut, hole, board, sc.dealer, R, sig,
                                                        up, t_weight=tw, tid=5, root=root)
        assert len(log) == len(sc.nodes), f"{sc.name}: {len(log)} decisions"
        for i, (nd, d, s) in enumerate(zip(sc.nodes, log, slots)):
            got = (int(d["player"]), int(d["street"]), int(d["history"]), legal_of(int(d["legal"])),
                   int(d["action"]), int(d["chips"]), int(d["pot"]), int(d["current_bet"]), int(d["lfr"]),
                   int(d["raises"]), queue_of(d))
            exp = (nd.player, nd.street, nd.history, nd.legal, nd.taken, nd.chips, nd.pot, nd.current_bet, nd.lfr,
                   nd.raises, nd.queue)
            assert got == exp, f"{sc.name} node {i}: got {got}, expected {exp}"
            assert int(d["slot"]) == s
            assert int(d["is_update"]) == (nd.player == up)
        assert int(tr["terminal"]) == sc.terminal
        assert int(tr["decisions"]) == len(sc.nodes)
        assert float(tr["utility"]) == u[up]
        upd = [(nd, s) for nd, s in zip(sc.nodes, slots) if nd.player == up]
        assert int(tr["updates"]) == len(upd)
        assert int(tr["update_legal"]) == sum(len(nd.legal) for nd, _ in upd)
        assert int(tr["path_hash"]) == fp.path_hash(slots, taken)
        # update-node contributions over all-zero regrets (cfr_gpu.cu:784-837):
        # sigma = 1/k over the k legal actions, q = max(1/k, 0.05), the sampled
        # action gets u/q - u and every other legal action 0 - u; strategy sums
        # get t * sigma.  Every node on the single path sees the terminal u.
        uf = np.float32(u[up])
        exp_c = []
        for nd, s in upd:
            k = len(nd.legal)
            sg = np.float32(1.0) / np.float32(k)
            q = max(sg, np.float32(0.05))
            for a in nd.legal:
                rd = np.float32(np.float32(uf / q) - uf) if a == nd.taken else np.float32(np.float32(0.0) - uf)
                exp_c.append((s, a, float(rd), float(np.float32(np.float32(tw) * sg))))
        got_c = sorted(zip(cs.tolist(), ca.tolist(), crd.tolist(), csd.tolist()))
        assert got_c == sorted(exp_c)


def test_forced_scenarios_cover_the_rules():
    """The scenarios exercise what the rules distinguish: every raise kind,
    a raise cap on both capped menus, a short all-in, all-in calls, folds,
    passive fallbacks from an illegal forced action, and all three terminals."""
    kinds = set()
    for sc in fp.SCENARIOS:
        cfg = sc.make_cfg(LOG2)
        for nd in sc.nodes:
            kinds.add((cfg.num_actions, cfg.menu[nd.taken].kind))
    for k in (mccfr.FOLD, mccfr.CHECK, mccfr.CALL, mccfr.RAISE_MIN, mccfr.POT_FRAC, mccfr.ALL_IN):
        assert (8, k) in kinds
    assert (4, mccfr.POT_FRAC) in kinds and (10, mccfr.POT_FRAC) in kinds
    assert {sc.terminal for sc in fp.SCENARIOS} == {1, 2, 3}
    fallbacks = [nd for sc in fp.SCENARIOS for nd in sc.nodes if nd.forced is not None and nd.forced != nd.taken]
    assert len(fallbacks) >= 3


# ---------------------------------------------------------------------------
# Two players with stacks at most the big blind: every hand is all-in from
# the blinds (SCENARIOS cover the betting; these cover the deal-to-showdown
# path of whole launches, against the brute-force evaluator).
# ---------------------------------------------------------------------------
def bf_ranks(hole, board):
    out = []
    for p in range(2):
        c = np.ascontiguousarray(np.concatenate([hole[2 * p: 2 * p + 2], board]), np.uint8)
        out.append(oracle.lib().orc_bf_class(oracle.p(c), 7))
    return out


def test_two_players_all_in_from_the_blinds(hand_table, golden_lut):
    """Stacks of 10 at blinds 10/20: both blinds are all-in (10 each), nobody
    acts, the hand runs to showdown: +10 / 0 / -10 by the brute-force ranks of
    the dealt cards.  No decision, so no contribution."""
    cfg = mccfr.reference_cfg(num_players=2, starting_stack=10, table_log2=12, seed=31)
    R = np.abs(np.random.default_rng(1).standard_normal((1 << 12) * 8)).astype(np.float32)
    n = 3000
    for up in (0, 1):
        dR, dS, aR, aS, tr = oracle.replay_launch(cfg, hand_table, golden_lut, R, None, up, 1, 0, n)
        assert (tr["decisions"] == 0).all() and (tr["terminal"] == 2).all()
        assert not aR.any() and not aS.any()
        for i in range(n):
            hole, board = oracle.deal_root(cfg, None, i)
            r = bf_ranks(hole, board)
            mine, other = r[up], r[1 - up]
            exp = 10.0 if mine > other else (-10.0 if mine < other else 0.0)
            assert float(tr["utility"][i]) == exp


def test_two_players_stack_of_one_big_blind(hand_table, golden_lut):
    """Stacks of 20: the big blind is all-in with its blind, the small blind
    (the dealer) faces 10 more with 10 behind: {fold, all-in} only.  A fold
    loses 10 (the big blind wins the 30 pot net +10); an all-in goes to
    showdown for +-20 or 0 by the brute-force ranks.  The update node of the
    small blind holds exactly those two legal actions."""
    cfg = mccfr.reference_cfg(num_players=2, starting_stack=20, table_log2=12, seed=32)
    R = np.abs(np.random.default_rng(2).standard_normal((1 << 12) * 8)).astype(np.float32)
    n = 3000
    seen = set()
    for i in range(n):
        hole, board = oracle.deal_root(cfg, None, i)
        cards = np.zeros(9, np.uint8)
        d = oracle.C.c_int()
        oracle.lib().orc_deal(cfg.seed, i, 0, 2, oracle.p(cards), oracle.C.byref(d))
        sb = d.value  # heads-up: the dealer posts the small blind
        r = bf_ranks(hole, board)
        for up in (0, 1):
            tr, log, (cs, ca, crd, csd) = oracle.walk_cards(cfg, hand_table, golden_lut, hole, board, sb, R, None,
                                                            up, tid=i)
            assert len(log) == 1
            d0 = log[0]
            assert int(d0["player"]) == sb and legal_of(int(d0["legal"])) == (0, 7)
            a = int(d0["action"])
            seen.add(a)
            if a == 0:
                exp = -10.0 if up == sb else 10.0
                assert int(tr["terminal"]) == 1
            else:
                mine, other = r[up], r[1 - up]
                exp = 20.0 if mine > other else (-20.0 if mine < other else 0.0)
                assert int(tr["terminal"]) == 2
            assert float(tr["utility"]) == exp
            assert len(cs) == (2 if up == sb else 0)
            if up == sb:
                assert sorted(ca.tolist()) == [0, 7]
    assert seen == {0, 7}
