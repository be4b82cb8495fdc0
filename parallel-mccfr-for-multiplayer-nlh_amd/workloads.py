"""Synthetic workloads of the batched and served forms.

* ``uneven_mix``: games with 2-6 players (uniform) and stacks log-uniform over
  10-200 bb, with the 6-action menu and <= 3 raises per street; a launch's
  traversals are split over the games by drawing each traversal's game.
* ``resolve_requests``: subgame re-solve requests with Poisson arrivals: street
  uniform over flop/turn/river, distinct board cards, 2-6 active seats, pot and
  stacks log-uniform over 5-200 bb, iteration budgets log-uniform over 1e3-1e5,
  table sizes log-uniform over 2^16-2^22 slots.

Chips use blinds 10/20 (1 bb = 20 chips).  Both generators are deterministic in
their seeds.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

import numpy as np

import mccfr

BB = 20


def _log_uniform(rng, lo, hi):
    return math.exp(rng.uniform(math.log(lo), math.log(hi)))


@dataclass
class Game:
    cfg: mccfr.GameCfg
    num_players: int
    stack: int


def uneven_mix(seed: int = 2, n_games: int = 32, table_log2: int = 20, traversals: int = 1 << 21):
    """(games, counts): n_games game configurations and the number of each
    launch's traversals that go to each game (multinomial over the games)."""
    rng = np.random.default_rng(seed)
    games: List[Game] = []
    for g in range(n_games):
        n = int(rng.integers(2, 7))
        stack = int(round(BB * _log_uniform(rng, 10.0, 200.0)))
        cfg = mccfr.menu_cfg(n, stack, mccfr.MENU_6, max_raises=3, table_log2=table_log2, seed=1000 + g)
        games.append(Game(cfg, n, stack))
    counts = rng.multinomial(traversals, np.full(n_games, 1.0 / n_games)).astype(np.int64)
    return games, counts


@dataclass
class Request:
    arrival: float            # seconds after the stream starts
    cfg: mccfr.GameCfg
    root: mccfr.RootState
    iterations: int
    batch: int
    street: int
    active: int


def request_batch(iterations: int) -> int:
    """Traversals per launch of a re-solve: about 32 rounds, 256..4096 lanes."""
    b = 1 << max(0, math.ceil(math.log2(max(1, math.ceil(iterations / 32)))))
    return int(min(4096, max(256, b)))


def random_root(rng, table_log2=None, seed=0):
    """(cfg, root) of one random public state."""
    n_active = int(rng.integers(2, 7))
    n_seats = int(rng.integers(n_active, 7))
    street = int(rng.integers(1, 4))
    seats = list(range(n_seats))
    active = sorted(rng.choice(seats, size=n_active, replace=False).tolist())
    pot = int(round(BB * _log_uniform(rng, 5.0, 200.0)))
    invested = [0] * 6
    stacks = [0] * 6
    folded_total = 0
    for p in seats:
        if p not in active:
            invested[p] = int(rng.integers(0, BB + 1))
            folded_total += invested[p]
    share = max(1, (pot - folded_total) // n_active)
    for p in active:
        invested[p] = share
        stacks[p] = int(round(BB * _log_uniform(rng, 5.0, 200.0)))
    for p in seats:
        if p not in active:
            stacks[p] = int(round(BB * _log_uniform(rng, 5.0, 200.0)))
    board = rng.choice(52, size=5, replace=False).tolist()
    dealer = int(rng.integers(0, n_seats))
    if table_log2 is None:
        table_log2 = int(rng.integers(16, 23))
    cfg = mccfr.menu_cfg(n_seats, 0, mccfr.MENU_6, max_raises=3, table_log2=table_log2, seed=seed)
    root = mccfr.root_state(street, dealer, active, stacks, invested, board)
    return cfg, root


def resolve_requests(n: int, rate: float = 200.0, arrival_seed: int = 3, state_seed: int = 4,
                     min_iters: float = 1e3, max_iters: float = 1e5) -> List[Request]:
    """n re-solve requests: Poisson arrivals at `rate` per second (seed
    arrival_seed), public states and budgets from seed state_seed."""
    ra = np.random.default_rng(arrival_seed)
    rs = np.random.default_rng(state_seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(ra.exponential(1.0 / rate))
        cfg, root = random_root(rs, seed=10_000 + i)
        iters = int(round(_log_uniform(rs, min_iters, max_iters)))
        out.append(Request(t, cfg, root, iters, request_batch(iters), int(root.street),
                           bin(int(root.active_mask)).count("1")))
    return out
