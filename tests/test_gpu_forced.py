"""The hand-worked forced-play sequences of tests/forced_play.py through the
GPU path (C ABI, one traversal per launch): for dealt hands whose dealer and
showdown outcome fit a scenario, one-hot regret rows at the scenario's slots
(so the device's BF16 snapshot is one-hot there) force every scripted node.
The trace must show the scenario's decisions, terminal, path hash and the
hand-worked utility; the tables must move by the scenario's contributions
(cfr_gpu.cu:784-837, CFR+ clamp), within the stated tolerance."""
import ctypes as C

import numpy as np
import pytest

import forced_play as fp
import helpers
import mccfr
import oracle

pytestmark = pytest.mark.gpu
LOG2 = 16
PER_SCENARIO = 4


def deals(sc, cfg, root, lut, hr, row):
    """First PER_SCENARIO traversal ids whose deal fits the scenario: the
    scenario's dealer (preflop), a showdown outcome it lists, distinct rows."""
    n = cfg.num_players
    out = []
    cards = np.zeros(2 * n + 5, np.uint8)
    d = C.c_int()
    for tid in range(20000):
        if root is None:
            oracle.lib().orc_deal(cfg.seed, tid, 0, n, oracle.p(cards), C.byref(d))
            if d.value != sc.dealer:
                continue
            hole, board = cards[: 2 * n].copy(), cards[2 * n:].copy()
        else:
            hole, board = oracle.deal_root(cfg, root, tid)
        u = fp.utilities(sc, fp.ranks_of(hole, board, oracle))
        if u is None:
            continue
        slots = fp.node_slots(sc, hole, board, lut, hr, LOG2, oracle)
        need = fp.forced_rows(sc, slots, row)
        if need is None:
            continue
        out.append((tid, slots, need, u))
        if len(out) == PER_SCENARIO:
            break
    assert len(out) == PER_SCENARIO
    return out


@pytest.mark.parametrize("sc", fp.SCENARIOS, ids=lambda s: s.name)
def test_forced_sequences_on_the_device(sc, hand_table, golden_lut):
    cfg = sc.make_cfg(LOG2)
    row = 8 if cfg.num_actions <= 8 else 16
    size = (1 << LOG2) * row
    taken = [nd.taken for nd in sc.nodes]
    tw = 2
    with mccfr.Context(cfg, 256) as ctx:
        ctx.set_hand_table(hand_table)
        root = fp.root_of(sc, cfg, fp.cards(*sc.board))
        tenant = 0 if root is None else ctx.tenant_create(cfg, root)
        for i, (tid, slots, need, u) in enumerate(deals(sc, cfg, root, golden_lut, hand_table, row)):
            up = i % cfg.num_players
            R0 = np.zeros(size, np.float32)
            for s, a in need.items():
                if a is not None:
                    R0[s * row + a] = 1.0
            S0 = np.zeros(size, np.float32)
            ctx.tenant_upload(tenant, R0, S0)
            print(f"forced {sc.name}: tid {tid} update player {up}", flush=True)
            tr = ctx.run_segments([(tenant, 1, up, tw, tid)], trace=True)[0]
            R1, S1, sig1 = ctx.tenant_download(tenant)
            assert int(tr["decisions"]) == len(sc.nodes)
            assert int(tr["terminal"]) == sc.terminal
            assert float(tr["utility"]) == u[up]
            assert int(tr["path_hash"]) == fp.path_hash(slots, taken)
            upd = [(nd, s) for nd, s in zip(sc.nodes, slots) if nd.player == up]
            assert int(tr["updates"]) == len(upd)
            assert int(tr["update_legal"]) == sum(len(nd.legal) for nd, _ in upd)
            # contributions: a node forced to a legal action has a one-hot
            # regret row, so sigma is one-hot and q = 1 (u/1 - u = 0 for the
            # action, -u for the other legal ones, t for its strategy sum);
            # an unforced node, or one forced to an illegal action, has sigma
            # uniform over its legal actions
            uf = np.float32(u[up])
            dR = np.zeros(size)
            dS = np.zeros(size)
            aR = np.zeros(size)
            for nd, s in upd:
                k = len(nd.legal)
                one_hot = nd.forced is not None and nd.forced == nd.taken
                for a in nd.legal:
                    sg = (np.float32(a == nd.taken) if one_hot else np.float32(1.0) / np.float32(k))
                    q = np.float32(1.0) if one_hot else max(np.float32(1.0) / np.float32(k), np.float32(0.05))
                    rd = np.float32(np.float32(uf / q) - uf) if a == nd.taken else np.float32(np.float32(0.0) - uf)
                    dR[s * row + a] += float(rd)
                    aR[s * row + a] += abs(float(rd))
                    dS[s * row + a] += float(np.float32(np.float32(tw) * sg))
            eR, eS = helpers.expected_post(R0, S0, dR, dS, cfr_plus=bool(cfg.use_cfr_plus))
            assert helpers.within_tol(R1, eR, aR).all()
            assert helpers.within_tol(S1, eS, np.abs(dS)).all()
            assert np.array_equal(sig1, oracle.regret_matching(R1, row))
