"""Hand-worked betting sequences for the parity tests of the walk under raises,
re-raises, all-ins and raise caps (cfr_gpu.cu:186-244 rules, :543-585
apply_action_gpu / rebuild_queue, :644-860 traverse_gpu).

Each scenario lists every decision node of one walk with the values worked out
by hand from the reference's rules, written as literals: who acts, on which
street, with which history bits, the legal actions, the action forced on the
node (a one-hot strategy row at the node's info-set slot; None = an all-zero
row, i.e. the passive fallback), the action taken (differs from the forced
one when that is illegal), and the state the action leaves: chips put in,
pot, bet to match, last full raise, raises this street and the seats still to
act in order.  The slot of a node is FNV-1a over the literal bytes (player,
hole bucket, board bucket, street, history), so a walk whose rules differ from
these sequences reaches other slots, falls back to passive play, and its log
no longer matches.

`cases` gives the terminal utility of every seat for each showdown outcome
(a predicate over the seats' 7-card ranks) -- also worked out by hand from the
side-pot rule (cfr_gpu.cu:694-752).
"""
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

import helpers
import mccfr

RANKS = "23456789TJQKA"
SUITS = "cdhs"  # suit-major codes: card = suit * 13 + rank


def card(s: str) -> int:
    return SUITS.index(s[1]) * 13 + RANKS.index(s[0])


def cards(*names: str) -> List[int]:
    return [card(n) for n in names]


@dataclass
class Node:
    player: int
    street: int
    history: int
    legal: Tuple[int, ...]
    forced: Optional[int]
    taken: int
    chips: int
    pot: int
    current_bet: int
    lfr: int
    raises: int
    queue: Tuple[int, ...]


@dataclass
class Scenario:
    name: str
    make_cfg: Callable[[int], mccfr.GameCfg]
    dealer: int
    nodes: List[Node]
    terminal: int  # 1 fold, 2 showdown, 3 depth limit
    cases: List[Tuple[Optional[Callable], Dict[int, float]]]
    hole: List[str] = field(default_factory=list)  # cards of the CPU walk, seat by seat
    board: List[str] = field(default_factory=list)
    root: Optional[dict] = None  # postflop root: street, stacks, invested (all seats active)


def N(player, street, history, legal, forced, taken, chips, pot, cb, lfr, raises, queue):
    return Node(player, street, history, tuple(legal), forced, taken, chips, pot, cb, lfr, raises, tuple(queue))


REF = lambda n, stack: (lambda log2: mccfr.reference_cfg(num_players=n, starting_stack=stack, table_log2=log2,
                                                         seed=77))
ALL8_FACING = (0, 2, 3, 4, 5, 6, 7)  # fold, call, min, 1/2, 2/3, pot, all-in
ALL8_OPEN = (1, 3, 4, 5, 6, 7)  # check, min, 1/2, 2/3, pot, all-in

# Reference menu indices: 0 F, 1 CHECK, 2 CALL, 3 MIN, 4 1/2 pot, 5 2/3 pot, 6 pot, 7 all-in.
SCENARIOS = [
    # Open pot, re-raise pot, fold, call; then check down.  After the re-raise
    # the queue is the seats after the raiser in seat order: the big blind
    # (seat 2) acts before the opener (seat 0).
    Scenario(
        "open_reraise_fold_call", REF(3, 1000), 0,
        [N(0, 0, 0, ALL8_FACING, 6, 6, 50, 80, 50, 30, 1, (1, 2)),
         N(1, 0, 0o6, ALL8_FACING, 6, 6, 120, 200, 130, 80, 2, (2, 0)),
         N(2, 0, 0o66, ALL8_FACING, 0, 0, 0, 200, 130, 80, 2, (0,)),
         N(0, 0, 0o660, ALL8_FACING, 2, 2, 80, 280, 130, 80, 2, ()),
         N(1, 1, 0o6602, ALL8_OPEN, None, 1, 0, 280, 0, 20, 0, (0,)),
         N(0, 1, 0o66021, ALL8_OPEN, None, 1, 0, 280, 0, 20, 0, ()),
         N(1, 2, 0o660211, ALL8_OPEN, None, 1, 0, 280, 0, 20, 0, (0,)),
         N(0, 2, 0o6602111, ALL8_OPEN, None, 1, 0, 280, 0, 20, 0, ()),
         N(1, 3, 0o66021111, ALL8_OPEN, None, 1, 0, 280, 0, 20, 0, (0,)),
         N(0, 3, 0o660211111, ALL8_OPEN, None, 1, 0, 280, 0, 20, 0, ())],
        2,
        [(lambda r: r[0] > r[1], {0: 150.0, 1: -130.0, 2: -20.0}),
         (lambda r: r[1] > r[0], {0: -130.0, 1: 150.0, 2: -20.0}),
         (lambda r: r[0] == r[1], {0: 10.0, 1: 10.0, 2: -20.0})],
        hole=["2h", "2s", "Ah", "Ad", "3c", "4d"], board=["2c", "2d", "7h", "9s", "Kc"]),
    # Flop root with a short stack: a half-pot bet of 150, a short all-in of
    # 250 (a raise of 100 < 150 keeps the last full raise at 150), then a
    # min-raise to 400 = 250 + 150, a call, and a check-down; the all-in seat
    # wins the main pot, the opener the side pot.
    Scenario(
        "short_allin_keeps_min_raise", REF(3, 1000), 2,
        [N(0, 1, 0, ALL8_OPEN, 4, 4, 150, 450, 150, 150, 1, (1, 2)),
         N(1, 1, 0o4, (0, 2, 7), 7, 7, 250, 700, 250, 150, 2, (2, 0)),
         N(2, 1, 0o47, ALL8_FACING, 3, 3, 400, 1100, 400, 150, 3, (0,)),
         N(0, 1, 0o473, (0, 2, 3, 4, 7), 2, 2, 250, 1350, 400, 150, 3, ()),
         N(0, 2, 0o4732, (1, 3, 7), None, 1, 0, 1350, 0, 20, 0, (2,)),
         N(2, 2, 0o47321, (1, 3, 7), None, 1, 0, 1350, 0, 20, 0, ()),
         N(0, 3, 0o473211, (1, 3, 7), None, 1, 0, 1350, 0, 20, 0, (2,)),
         N(2, 3, 0o4732111, (1, 3, 7), None, 1, 0, 1350, 0, 20, 0, ())],
        2,
        [(lambda r: r[1] > r[0] > r[2], {0: -200.0, 1: 700.0, 2: -500.0}),
         (lambda r: r[1] > r[2] > r[0], {0: -500.0, 1: 700.0, 2: -200.0}),
         (lambda r: r[0] > r[1] and r[0] > r[2], {0: 850.0, 1: -350.0, 2: -500.0}),
         (lambda r: r[2] > r[1] and r[2] > r[0], {0: -500.0, 1: -350.0, 2: 850.0})],
        hole=["Ac", "Ks", "Kc", "Kd", "5s", "6s"], board=["Kh", "7d", "2c", "9s", "3h"],
        root=dict(street=1, stacks=[1000, 250, 1000], invested=[100, 100, 100])),
    # {F, C/C, pot, all-in}, at most 2 raises per street: after two raises the
    # big blind and the opener may only fold or call (a forced raise or all-in
    # falls back to the call); on the flop a bet, an all-in raise, a fold to a
    # bet covering the stack ({F, all-in} only), an all-in call, and the hand
    # goes to showdown with nobody left to act.
    Scenario(
        "raise_cap_four_action_menu",
        lambda log2: mccfr.menu_cfg(3, 1000, mccfr.MENU_FCPA, max_raises=2, table_log2=log2, seed=77), 0,
        [N(0, 0, 0, (0, 1, 2, 3), 2, 2, 50, 80, 50, 30, 1, (1, 2)),
         N(1, 0, 0b10, (0, 1, 2, 3), 2, 2, 120, 200, 130, 80, 2, (2, 0)),
         N(2, 0, 0b1010, (0, 1), 2, 1, 110, 310, 130, 80, 2, (0,)),
         N(0, 0, 0b101001, (0, 1), 3, 1, 80, 390, 130, 80, 2, ()),
         N(1, 1, 0b10100101, (1, 2, 3), 2, 2, 390, 780, 390, 390, 1, (2, 0)),
         N(2, 1, 0b1010010110, (0, 1, 3), 3, 3, 870, 1650, 870, 480, 2, (0, 1)),
         N(0, 1, 0b101001011011, (0, 3), None, 0, 0, 1650, 870, 480, 2, (1,)),
         N(1, 1, 0b10100101101100, (0, 3), 3, 3, 480, 2130, 870, 480, 2, ())],
        2,
        [(lambda r: r[2] > r[1], {0: -130.0, 1: -1000.0, 2: 1130.0}),
         (lambda r: r[1] > r[2], {0: -130.0, 1: 1130.0, 2: -1000.0}),
         (lambda r: r[1] == r[2], {0: -130.0, 1: 65.0, 2: 65.0})],
        hole=["3h", "5s", "Kc", "Kd", "Ac", "Ad"], board=["2c", "7d", "9h", "Js", "4c"]),
    # Ten-action menu {F, C/C, 1/4, 1/2, 3/4, 1, 3/2, 2, 3 pot, all-in}, 4-bit
    # history, at most 4 raises per street: a 3/4-pot open (to 20 + 90 / 4 =
    # 42), 1/2 pot, 3 pot, 3/2 pot (the fourth raise), then the cap leaves
    # fold/call; on the flop only the raises that fit the stack are legal,
    # all-in and an all-in call.
    Scenario(
        "raise_cap_ten_action_menu",
        lambda log2: mccfr.menu_cfg(3, 4000, mccfr.MENU_10, max_raises=4, table_log2=log2, seed=77), 0,
        [N(0, 0, 0x0, range(10), 4, 4, 42, 72, 42, 22, 1, (1, 2)),
         N(1, 0, 0x4, range(10), 3, 3, 68, 140, 78, 36, 2, (2, 0)),
         N(2, 0, 0x43, range(10), 8, 8, 478, 618, 498, 420, 3, (0, 1)),
         N(0, 0, 0x438, range(10), 6, 6, 1383, 2001, 1425, 927, 4, (1, 2)),
         N(1, 0, 0x4386, (0, 1), 7, 1, 1347, 3348, 1425, 927, 4, (2,)),
         N(2, 0, 0x43861, (0, 1), 0, 0, 0, 3348, 1425, 927, 4, ()),
         N(1, 1, 0x438610, (1, 2, 3, 4, 9), 9, 9, 2575, 5923, 2575, 2575, 1, (0,)),
         N(0, 1, 0x4386109, (0, 9), 9, 9, 2575, 8498, 2575, 2575, 1, ())],
        2,
        [(lambda r: r[0] == r[1], {0: 249.0, 1: 249.0, 2: -498.0}),
         (lambda r: r[0] > r[1], {0: 4498.0, 1: -4000.0, 2: -498.0}),
         (lambda r: r[1] > r[0], {0: -4000.0, 1: 4498.0, 2: -498.0})],
        hole=["2c", "3c", "4d", "5d", "6h", "7h"], board=["As", "Ks", "Qs", "Js", "Ts"]),
    # River root with stacks 300 / 600 / 900 over 100 each: three all-ins make
    # three pots (1,200 for all, 600 for seats 1-2, 300 returned to seat 2).
    Scenario(
        "three_way_allin_side_pots", REF(3, 1000), 2,
        [N(0, 3, 0, ALL8_OPEN, 7, 7, 300, 600, 300, 300, 1, (1, 2)),
         N(1, 3, 0o7, (0, 2, 3, 4, 7), 7, 7, 600, 1200, 600, 300, 2, (2,)),
         N(2, 3, 0o77, (0, 2, 3, 7), 7, 7, 900, 2100, 900, 300, 3, ())],
        2,
        [(lambda r: r[0] > r[1] > r[2], {0: 800.0, 1: -100.0, 2: -700.0}),
         (lambda r: r[0] > r[2] > r[1], {0: 800.0, 1: -700.0, 2: -100.0}),
         (lambda r: r[1] > r[0] and r[1] > r[2], {0: -400.0, 1: 1100.0, 2: -700.0}),
         (lambda r: r[2] > r[0] and r[2] > r[1], {0: -400.0, 1: -700.0, 2: 1100.0})],
        hole=["Ac", "Ad", "Kc", "Kd", "3h", "5s"], board=["2c", "7d", "9h", "Js", "4c"],
        root=dict(street=3, stacks=[300, 600, 900], invested=[100, 100, 100])),
    # Heads-up: the dealer posts the small blind and acts first preflop, the
    # big blind last (its option: a min-raise to 40); after the flop the big
    # blind acts first; a half-pot bet and a fold.
    Scenario(
        "heads_up_order_and_option", REF(2, 1000), 0,
        [N(0, 0, 0, ALL8_FACING, 2, 2, 10, 40, 20, 20, 0, (1,)),
         N(1, 0, 0o2, ALL8_OPEN, 3, 3, 20, 60, 40, 20, 1, (0,)),
         N(0, 0, 0o23, ALL8_FACING, 2, 2, 20, 80, 40, 20, 1, ()),
         N(1, 1, 0o232, ALL8_OPEN, None, 1, 0, 80, 0, 20, 0, (0,)),
         N(0, 1, 0o2321, ALL8_OPEN, 4, 4, 40, 120, 40, 40, 1, (1,)),
         N(1, 1, 0o23214, ALL8_FACING, 0, 0, 0, 120, 40, 40, 1, ())],
        1, [(None, {0: 40.0, 1: -40.0})],
        hole=["Ah", "Kh", "2c", "7d"], board=["3s", "8h", "9c", "Td", "Jd"]),
    # Four players fold to the big blind, which wins the blinds without acting.
    Scenario(
        "fold_around", REF(4, 1000), 0,
        [N(3, 0, 0, ALL8_FACING, 0, 0, 0, 30, 20, 20, 0, (0, 1, 2)),
         N(0, 0, 0, ALL8_FACING, 0, 0, 0, 30, 20, 20, 0, (1, 2)),
         N(1, 0, 0, ALL8_FACING, 0, 0, 0, 30, 20, 20, 0, (2,))],
        1, [(None, {0: 0.0, 1: -10.0, 2: 10.0, 3: 0.0})],
        hole=["Ah", "Kh", "2c", "7d", "Qs", "Qd", "5c", "6c"], board=["3s", "8h", "9c", "Td", "Jd"]),
    # Heads-up min-raise war: 20 decisions, the history keeps the last 10
    # actions (30 bits) from the 11th node on, and the depth guard of 20 ends
    # the walk with utility 0.
    Scenario(
        "min_raise_war_depth_guard", REF(2, 10000), 0,
        [N(0, 0, 0, ALL8_FACING, 3, 3, 30, 60, 40, 20, 1, (1,))]
        + [N(k % 2, 0, int("3" * min(k, 10), 8), ALL8_FACING, 3, 3, 40, 60 + 40 * k, 40 + 20 * k, 20, k + 1,
             ((k + 1) % 2,)) for k in range(1, 20)],
        3, [(None, {0: 0.0, 1: 0.0})],
        hole=["Ah", "Kh", "2c", "7d"], board=["3s", "8h", "9c", "Td", "Jd"]),
    # Three-way limp, the big blind's option raise (pot, to 80) reopens the
    # betting for both limpers; one folds, one calls, check-down.
    Scenario(
        "big_blind_option_raise", REF(3, 1000), 0,
        [N(0, 0, 0, ALL8_FACING, 2, 2, 20, 50, 20, 20, 0, (1, 2)),
         N(1, 0, 0o2, ALL8_FACING, 2, 2, 10, 60, 20, 20, 0, (2,)),
         N(2, 0, 0o22, ALL8_OPEN, 6, 6, 60, 120, 80, 60, 1, (0, 1)),
         N(0, 0, 0o226, ALL8_FACING, 0, 0, 0, 120, 80, 60, 1, (1,)),
         N(1, 0, 0o2260, ALL8_FACING, 2, 2, 60, 180, 80, 60, 1, ()),
         N(1, 1, 0o22602, ALL8_OPEN, None, 1, 0, 180, 0, 20, 0, (2,)),
         N(2, 1, 0o226021, ALL8_OPEN, None, 1, 0, 180, 0, 20, 0, ()),
         N(1, 2, 0o2260211, ALL8_OPEN, None, 1, 0, 180, 0, 20, 0, (2,)),
         N(2, 2, 0o22602111, ALL8_OPEN, None, 1, 0, 180, 0, 20, 0, ()),
         N(1, 3, 0o226021111, ALL8_OPEN, None, 1, 0, 180, 0, 20, 0, (2,)),
         N(2, 3, 0o2260211111, ALL8_OPEN, None, 1, 0, 180, 0, 20, 0, ())],
        2,
        [(lambda r: r[1] > r[2], {0: -20.0, 1: 100.0, 2: -80.0}),
         (lambda r: r[2] > r[1], {0: -20.0, 1: -80.0, 2: 100.0}),
         (lambda r: r[1] == r[2], {0: -20.0, 1: 10.0, 2: 10.0})],
        hole=["5d", "6d", "Qc", "Qd", "3h", "3s"], board=["2c", "7d", "9h", "Js", "4c"]),
]


def by_name(name: str) -> Scenario:
    return next(s for s in SCENARIOS if s.name == name)


def root_of(sc: Scenario, cfg, board_cards):
    """The scenario's postflop root (None for the preflop root)."""
    if sc.root is None:
        return None
    n = cfg.num_players
    st = sc.root["street"]
    return mccfr.root_state(st, sc.dealer, (1 << n) - 1, sc.root["stacks"], sc.root["invested"],
                            list(board_cards[: st + 2]))


def slot_of(player, hb, bkt, street, history, log2):
    """Info-set slot: FNV-1a over (player, hole bucket, board bucket, street,
    history bytes little-endian) masked to log2 bits (cfr_gpu.cu:165-179)."""
    data = bytes([player, hb, bkt, street]) + int(history).to_bytes(4, "little")
    return helpers.fnv1a(data) & ((1 << log2) - 1)


def node_slots(sc: Scenario, hole, board, lut, hr, log2, orc):
    """Slot of every node of the scenario for the given deal."""
    out = []
    for nd in sc.nodes:
        h = np.ascontiguousarray(hole[2 * nd.player: 2 * nd.player + 2], np.uint8)
        hb = int(lut[int(h[0]) * 52 + int(h[1])])
        bkt = 0 if nd.street == 0 else orc.lib().orc_board_bucket(orc.p(hr), orc.p(h),
                                                                  orc.p(np.ascontiguousarray(board, np.uint8)),
                                                                  nd.street)
        out.append(slot_of(nd.player, hb, bkt, nd.street, nd.history, log2))
    return out


def forced_rows(sc: Scenario, slots, row):
    """{slot: action} of the scripted nodes, or None when two nodes that need
    different rows share a slot (a hash collision for this deal)."""
    need = {}
    for nd, s in zip(sc.nodes, slots):
        want = nd.forced
        if s in need and need[s] != want:
            return None
        need[s] = want
    return need


def ranks_of(hole, board, orc):
    n = len(hole) // 2
    out = []
    for p in range(n):
        c = np.ascontiguousarray(list(hole[2 * p: 2 * p + 2]) + list(board), np.uint8)
        out.append(orc.lib().orc_bf_class(orc.p(c), 7))
    return out


def utilities(sc: Scenario, ranks):
    """Hand-worked utility of every seat for these showdown ranks (None if no case applies)."""
    for pred, u in sc.cases:
        if pred is None or pred(ranks):
            return u
    return None


def path_hash(slots, actions):
    """FNV-1a over (slot bytes, action byte) of every decision, as the trace records it."""
    data = b"".join(int(s).to_bytes(4, "little") + bytes([a & 0xFF]) for s, a in zip(slots, actions))
    return helpers.fnv1a(data)
