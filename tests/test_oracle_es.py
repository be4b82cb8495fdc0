"""Pins the oracle's external-sampling mode (mccfr.h MCCFR_SAMPLE_EXTERNAL)
by brute force on a tiny game: heads-up, 60-chip stacks at blinds 10/20, the
reference's 8-action menu, where the whole tree is enumerated.

The brute force walks every action of both players from a dealt hand and
forms the exact counterfactual quantities external sampling estimates: at an
update-player node h, v(h, a) for every legal a (opponent nodes averaged with
their sampling probabilities: the BF16 snapshot's legal entries over their sum,
or the passive fallback with certainty), v(h) = sum_a sigma(a) v(h, a) with
sigma from the FP32 regrets, and the expected contributions
  E[regret delta(I, a)] = sum over h in I of p_opp(h) (v(h, a) - v(h)),
  E[strategy-sum delta(I, a)] = sum over h in I of p_opp(h) t sigma(a),
p_opp(h) the opponent's probability of reaching h.  With a deterministic
opponent (one-hot snapshot rows) one external-sampling traversal equals them;
with a mixed opponent the mean over many traversals of one hand (their draws
differ) converges to them.  The rules themselves come from the oracle's
pinned rule functions (orc_valid_mask, orc_chips: the reference's own
abstraction.cpp fixtures) and the info-set hash; the brute force's own parts
are the tree enumeration, the queue and street order of a heads-up hand, the
showdown and the averaging."""
import ctypes as C

import numpy as np
import pytest

import mccfr
import oracle

O = oracle.lib()
LOG2 = 12
ROW = 8


def cfg_tiny():
    cfg = mccfr.reference_cfg(num_players=2, starting_stack=60, table_log2=LOG2, seed=5)
    cfg.sampling = mccfr.SAMPLE_EXTERNAL
    return cfg


def rank7(cards):
    c = np.ascontiguousarray(cards, np.uint8)
    return O.orc_bf_class(oracle.p(c), 7)


class Brute:
    """Exact expected contributions of one hand (2 players, dealer = small blind)."""

    def __init__(self, cfg, hr, lut, hole, board, dealer, up, R, sig, tw):
        self.cfg, self.up, self.tw = cfg, up, tw
        self.R = R.reshape(-1, ROW)
        self.sig = sig.reshape(-1, ROW)
        self.hole, self.board = hole, board
        h = [np.ascontiguousarray(hole[2 * p: 2 * p + 2], np.uint8) for p in range(2)]
        self.hb = [int(lut[int(x[0]) * 52 + int(x[1])]) for x in h]
        self.bk = [[0] + [O.orc_board_bucket(oracle.p(hr), oracle.p(x), oracle.p(np.ascontiguousarray(board,
                                                                                                     np.uint8)), st)
                          for st in (1, 2, 3)] for x in h]
        self.ranks = [rank7(np.concatenate([hole[2 * p: 2 * p + 2], board])) for p in range(2)]
        self.dealer = dealer
        self.dR = {}
        self.dS = {}

    def bf16(self, x):
        return float(np.frombuffer(np.array([int(x) << 16], np.uint32).tobytes(), np.float32)[0])

    def run(self):
        d = self.dealer
        sbp, bbp = d, 1 - d
        st = dict(stacks=[60, 60], bets=[0, 0], inv=[0, 0], folded=[False, False], allin=[False, False],
                  pot=30, cb=20, lfr=20, street=0, ab=0, raises=0)
        st["stacks"][sbp], st["bets"][sbp], st["inv"][sbp] = 50, 10, 10
        st["stacks"][bbp], st["bets"][bbp], st["inv"][bbp] = 40, 20, 20
        st["queue"] = [sbp, bbp]  # heads-up: the small blind first, the big blind last
        return self.node(st, 1.0, 0)

    def payoff(self, st):
        up, other = self.up, 1 - self.up
        if st["folded"][up]:
            return -float(st["inv"][up])
        if st["folded"][other]:
            return float(st["pot"] - st["inv"][up])
        assert st["inv"][0] == st["inv"][1]  # equal stacks: equal investments at a showdown
        if self.ranks[up] > self.ranks[other]:
            return float(st["pot"] - st["inv"][up])
        if self.ranks[up] < self.ranks[other]:
            return -float(st["inv"][up])
        return float(st["pot"]) / 2.0 - st["inv"][up]

    def child(self, st, p, a, vm):
        s = {k: (list(v) if isinstance(v, list) else v) for k, v in st.items()}
        to_call = s["cb"] - s["bets"][p]
        chips = O.orc_chips(C.byref(self.cfg), a, s["pot"], s["stacks"][p], to_call, s["cb"], s["lfr"])
        s["ab"] = ((s["ab"] << 3) | a) & 0x3FFFFFFF
        s["queue"] = s["queue"][1:]
        kind = self.cfg.menu[a].kind
        if kind == mccfr.FOLD:
            s["folded"][p] = True
            return s
        if kind == mccfr.CHECK:
            return s
        s["stacks"][p] -= chips
        s["bets"][p] += chips
        s["inv"][p] += chips
        s["pot"] += chips
        if s["stacks"][p] == 0:
            s["allin"][p] = True
        if s["bets"][p] > s["cb"]:
            raise_ = s["bets"][p] - s["cb"]
            if raise_ >= s["lfr"] or s["cb"] == 0:
                s["lfr"] = raise_
            s["cb"] = s["bets"][p]
            s["raises"] += 1
            o = 1 - p
            s["queue"] = [o] if not (s["folded"][o] or s["allin"][o]) else []
        return s

    def node(self, st, reach, depth):
        assert depth < 20
        if any(st["folded"]):
            return self.payoff(st)
        if not st["queue"]:
            if st["street"] == 3:
                return self.payoff(st)
            s = {k: (list(v) if isinstance(v, list) else v) for k, v in st.items()}
            s["street"] += 1
            s["bets"], s["cb"], s["lfr"], s["raises"] = [0, 0], 0, 20, 0
            s["queue"] = [q for q in ((self.dealer + 1) % 2, self.dealer) if not (s["folded"][q] or s["allin"][q])]
            if not s["queue"]:
                return self.payoff(s)
            return self.node(s, reach, depth + 1)
        p = st["queue"][0]
        to_call = st["cb"] - st["bets"][p]
        vm = O.orc_valid_mask(C.byref(self.cfg), st["pot"], st["stacks"][p], to_call, st["cb"], st["lfr"],
                              st["raises"])
        legal = [a for a in range(8) if vm >> a & 1]
        slot = O.orc_info_hash(p, self.hb[p], self.bk[p][st["street"]], st["street"], st["ab"], LOG2)
        if p == self.up:
            r = [max(0.0, float(self.R[slot, a])) if a in legal else 0.0 for a in range(8)]
            tot = sum(r)
            sig = [(r[a] / tot if tot > 1e-7 else (1.0 / len(legal) if a in legal else 0.0)) for a in range(8)]
            v = {a: self.node(self.child(st, p, a, vm), reach, depth + 1) for a in legal}
            vi = sum(sig[a] * v[a] for a in legal)
            for a in legal:
                self.dR[(slot, a)] = self.dR.get((slot, a), 0.0) + reach * (v[a] - vi)
                self.dS[(slot, a)] = self.dS.get((slot, a), 0.0) + reach * self.tw * sig[a]
            return vi
        w = {a: self.bf16(self.sig[slot, a]) for a in legal}
        tot = sum(w.values())
        if tot < 1e-7:
            probs = {O.orc_passive_fallback(C.byref(self.cfg), vm): 1.0}
        else:
            probs = {a: w[a] / tot for a in legal if w[a] > 0}
        return sum(q * self.node(self.child(st, p, a, vm), reach * q, depth + 1) for a, q in probs.items())


def es_sums(cfg, hr, lut, hole, board, dealer, R, sig, up, tw, tids):
    """Mean over traversals (one per id, same hand) of the oracle's ES
    contributions, by (slot, action), and the mean root value."""
    dR, dS, u = {}, {}, 0.0
    for tid in tids:
        tr, _, (cs, ca, crd, csd) = oracle.walk_cards(cfg, hr, lut, hole, board, dealer, R, sig, up, t_weight=tw,
                                                      tid=tid, cap=4096)
        u += float(tr["utility"])
        for s_, a_, r_, d_ in zip(cs.tolist(), ca.tolist(), crd.tolist(), csd.tolist()):
            dR[(s_, a_)] = dR.get((s_, a_), 0.0) + r_
            dS[(s_, a_)] = dS.get((s_, a_), 0.0) + d_
    k = len(tids)
    return {x: y / k for x, y in dR.items()}, {x: y / k for x, y in dS.items()}, u / k


def hand(seed):
    rng = np.random.default_rng(seed)
    cards = rng.permutation(52)[:9].astype(np.uint8)
    return cards[:4], cards[4:]


@pytest.mark.parametrize("seed", range(6))
def test_es_equals_brute_force_with_a_deterministic_opponent(hand_table, golden_lut, seed):
    cfg = cfg_tiny()
    rng = np.random.default_rng(100 + seed)
    R = (rng.random((1 << LOG2) * ROW) * 2.0 - 0.5).astype(np.float32)  # mixed update-player strategies
    # one-hot snapshot rows on a random action (an illegal one falls back to passive play)
    sig = np.zeros((1 << LOG2) * ROW, np.uint16)
    sig.reshape(-1, ROW)[np.arange(1 << LOG2), rng.integers(0, 8, 1 << LOG2)] = 0x3F80
    hole, board = hand(seed)
    compared = 0
    for up in (0, 1):
        for dealer in (0, 1):
            bf = Brute(cfg, hand_table, golden_lut, hole, board, dealer, up, R, sig, 3.0)
            v_exact = bf.run()
            dR, dS, u = es_sums(cfg, hand_table, golden_lut, hole, board, dealer, R, sig, up, 3, [seed])
            assert abs(u - v_exact) <= 1e-4 * max(1.0, abs(v_exact))
            assert set(dR) == set(bf.dR)
            compared += len(dR)
            for key, x in bf.dR.items():
                assert abs(dR[key] - x) <= 1e-4 * max(1.0, abs(x)), (key, dR[key], x)
                assert abs(dS[key] - bf.dS[key]) <= 1e-5 * max(1.0, bf.dS[key])
    assert compared > 10


def test_es_is_unbiased_against_brute_force_with_a_mixed_opponent(hand_table, golden_lut):
    cfg = cfg_tiny()
    rng = np.random.default_rng(7)
    R = (rng.random((1 << LOG2) * ROW) * 2.0 - 0.5).astype(np.float32)
    sig = oracle.regret_matching(np.abs(R) + 0.05, ROW)  # every legal action possible for the opponent
    hole, board = hand(11)
    K = 3000
    checked = 0
    for up, dealer in ((0, 0), (1, 0)):
        bf = Brute(cfg, hand_table, golden_lut, hole, board, dealer, up, R, sig, 1.0)
        v_exact = bf.run()
        per = []
        for tid in range(K):
            per.append(es_sums(cfg, hand_table, golden_lut, hole, board, dealer, R, sig, up, 1, [tid]))
        us = np.array([p[2] for p in per])
        assert abs(us.mean() - v_exact) <= 5 * us.std() / np.sqrt(K) + 1e-3
        seen = set().union(*(p[0].keys() for p in per))
        assert seen <= set(bf.dR)  # the sampled trees lie inside the full one
        for key, x in bf.dR.items():
            visits = np.array([key in p[0] for p in per])
            if visits.mean() < 0.02:
                continue  # too rare for a mean over K traversals
            xs = np.array([p[0].get(key, 0.0) for p in per])
            se = xs.std() / np.sqrt(K)
            assert abs(xs.mean() - x) <= 5 * se + 1e-3 * max(1.0, abs(x)), (key, xs.mean(), x, se)
            ys = np.array([p[1].get(key, 0.0) for p in per])
            assert abs(ys.mean() - bf.dS[key]) <= 5 * ys.std() / np.sqrt(K) + 1e-6
            checked += 1
    assert checked > 20


def test_es_zero_sum_of_regrets_weighted_by_sigma(hand_table, golden_lut):
    """At every node sum_a sigma(a) (v(a) - v(I)) = 0, so for a slot reached
    once (one traversal) the regret deltas weighted by the strategy-sum deltas
    (t sigma, t = 1) add to zero, up to FP32 rounding."""
    cfg = mccfr.cfg_3p_20bb(table_log2=16, seed=2)
    cfg.sampling = mccfr.SAMPLE_EXTERNAL
    rng = np.random.default_rng(3)
    R = (rng.random((1 << 16) * ROW) * 2.0 - 0.5).astype(np.float32)
    sig = oracle.regret_matching(np.maximum(R, 0), ROW)
    n_checked = 0
    for tid in range(200):
        hole3 = np.ascontiguousarray(np.random.default_rng(tid).permutation(52)[:11], np.uint8)
        tr, log, (cs, ca, crd, csd) = oracle.walk_cards(cfg, hand_table, golden_lut, hole3[:6], hole3[6:], tid % 3,
                                                        R, sig, tid % 3, tid=tid, cap=4096)
        by = {}
        for s_, a_, r_, d_ in zip(cs.tolist(), ca.tolist(), crd.tolist(), csd.tolist()):
            by.setdefault(s_, []).append((r_, d_))
        for s_, xs in by.items():
            if len(xs) != len({a for a in ca[cs == s_].tolist()}):
                continue  # the slot was reached more than once
            tot = sum(r * d for r, d in xs)
            mag = sum(abs(r * d) for r, d in xs)
            # FP32: v(I) and each v(a) - v(I) round at ~6e-8 of |v| <= 800 chips
            assert abs(tot) <= 1e-5 * mag + 2e-4 * sum(d for _, d in xs)
            n_checked += 1
    assert n_checked > 100
