"""The identity the record emission (k_emit, traverse.hip) uses for the update player's
showdown payoff, checked bit-for-bit against the reference's side-pot loop
(cfr_gpu.cu:694-752: ascending investment levels, each layer shared by the
best live ranks that reach it, FP32 shares added in level order).

Without a live seat of the update player's rank investing more than M (the
largest investment of a live seat with a better rank), the update player wins
exactly the layers in (M, inv_up], alone, and their pots add up to
sum_p min(inv_p, inv_up) - min(inv_p, M); every share is a whole number of
chips below 2^24, so the loop's FP32 sum is exact.  Ties keep the loop.  The
GPU parity tests cover the kernel itself; this pins the algebra, including
zero investments (the sign of a zero payoff) and split pots."""
import random

import numpy as np

f32 = np.float32
SEATS = 6


def layer_loop(inv, rk, folded, up):
    """cfr_gpu.cu:694-752 restated for one update player."""
    inv_up, rk_up = inv[up], rk[up]
    up_in = not (folded >> up & 1)
    pay = f32(-f32(inv_up))
    prev = 0
    for _ in range(SEATS):
        above = [x for x in inv if x > prev]
        if not above:
            continue
        lvl = min(above)
        pot = sum(max(0, min(x, lvl) - prev) for x in inv)
        elig = [p for p in range(SEATS) if not (folded >> p & 1) and inv[p] >= lvl]
        best = max((rk[p] for p in elig), default=0)
        if best > 0:
            n_win = sum(1 for p in elig if rk[p] == best)
            if up_in and inv_up >= lvl and rk_up == best:
                pay = f32(pay + f32(f32(pot) / f32(n_win)))
        prev = lvl
    return pay


def closed_form(inv, rk, folded, up):
    """The kernel's form; None where it defers to the loop (a tie above M)."""
    inv_up, rk_up = inv[up], rk[up]
    up_in = not (folded >> up & 1)
    live = [not (folded >> p & 1) and p != up for p in range(SEATS)]
    m = max([inv[p] for p in range(SEATS) if live[p] and rk[p] > rk_up], default=0)
    if up_in and any(live[p] and rk[p] == rk_up and inv[p] > m for p in range(SEATS)):
        return None
    pay = f32(-f32(inv_up))
    if up_in and m < inv_up:
        won = sum(min(x, inv_up) - min(x, m) for x in inv)
        pay = f32(pay + f32(won))
    return pay


def random_state(rng):
    n = rng.randint(2, SEATS)
    inv = [rng.choice([0, 10, 20, 40, 100, 300, 2000, 4000, rng.randint(1, 32767)]) for _ in range(n)]
    rk = [rng.choice([rng.randint(1, 7462), 5000, 5000, 7000]) for _ in range(n)]
    folded = sum(1 << p for p in range(n) if rng.random() < 0.35)
    return inv + [0] * (SEATS - n), rk + [0] * (SEATS - n), folded, rng.randrange(n)


def test_closed_form_matches_layer_loop_bitwise():
    rng = random.Random(2026)
    used = ties = 0
    for _ in range(20000):
        inv, rk, folded, up = random_state(rng)
        c = closed_form(inv, rk, folded, up)
        if c is None:
            ties += 1
            continue
        used += 1
        assert c.tobytes() == layer_loop(inv, rk, folded, up).tobytes(), (inv, rk, folded, up)
    assert used > 15000 and ties > 100  # both branches exercised


def test_known_states():
    # three-way all-in: up (rank 7000, 4000 in) beats the 2000 stack but loses to nobody
    inv, rk = [4000, 2000, 4000, 0, 0, 0], [7000, 6000, 5000, 0, 0, 0]
    assert closed_form(inv, rk, 0, 0) == layer_loop(inv, rk, 0, 0) == f32(6000)
    # a better hand covers only the main pot: up wins the side pot alone
    inv, rk = [4000, 2000, 4000, 0, 0, 0], [5000, 7000, 4000, 0, 0, 0]
    assert closed_form(inv, rk, 0, 0) == layer_loop(inv, rk, 0, 0) == f32(0)
    # folded up player: loses its investment
    assert closed_form(inv, rk, 1, 0) == layer_loop(inv, rk, 1, 0) == f32(-4000)
    # nothing invested: -0.0 as in the loop
    z = closed_form([0] * 6, [5] * 2 + [0] * 4, 0, 0)
    assert z.tobytes() == layer_loop([0] * 6, [5] * 2 + [0] * 4, 0, 0).tobytes() == f32(-0.0).tobytes()
    # a split pot defers to the loop
    assert closed_form([100, 100, 0, 0, 0, 0], [6000, 6000, 0, 0, 0, 0], 0, 0) is None
