"""Games drawn at random, each checked launch by launch against the oracle.

The fixed configurations elsewhere cover the reference's menu and the
benchmark menus; these cover the rest of the configuration space the C ABI
accepts: menus of 3-16 actions built from every action kind (check and call
as separate entries or one check/call, min-raises, pot fractions with
numerators and denominators up to 255, all-in) in any order, 2-6 seats,
blinds from 2 to 100 chips, stacks from below one big blind to 32,767 chips,
raise caps 0 (unlimited) to 4, history widths up to 8 bits, CFR+ and linear
weighting on and off, preflop roots and random flop / turn / river roots
(active seats, stacks, prior investments, board), from zero and from random
tables, in both samplings.  For every launch: each traversal's record
(utility bits, path hash, decision and update counts, terminal kind)
bit-exact, the tables within the FP64 replay's tolerance, the BF16 snapshot
equal to regret matching of the new regrets.  Each case's parameters are
printed before it runs."""
import numpy as np
import pytest

import helpers
import mccfr
import oracle

pytestmark = pytest.mark.gpu
FIELDS = ("path_hash", "decisions", "updates", "terminal", "update_legal")


def random_menu(rng, n_actions):
    """fold, a passive entry (check/call, or check and call apart), all-in,
    and raises; shuffled."""
    menu = [(mccfr.FOLD, 0, 1), (mccfr.ALL_IN, 0, 1)]
    if rng.random() < 0.5:
        menu.append((mccfr.CHECK_CALL, 0, 1))
    else:
        menu += [(mccfr.CHECK, 0, 1), (mccfr.CALL, 0, 1)]
    while len(menu) < n_actions:
        u = rng.random()
        if u < 0.15:
            menu.append((mccfr.RAISE_MIN, 0, 1))
        elif u < 0.25:  # extreme fractions: 255 pots, or 1/255 of one (a min-raise in effect)
            menu.append((mccfr.POT_FRAC, 255, 1) if rng.random() < 0.5 else (mccfr.POT_FRAC, 1, 255))
        else:
            menu.append((mccfr.POT_FRAC, int(rng.integers(1, 9)), int(rng.integers(1, 9))))
    order = rng.permutation(len(menu))
    return [menu[i] for i in order]


def random_game(seed, external=False):
    rng = np.random.default_rng(seed)
    N = int(rng.integers(2, 7))
    bb = int(rng.choice([2, 10, 20, 50, 100]))
    sb = int(rng.integers(1, bb + 1))
    if external:  # trees that fit the scratch: shallow stacks, small menus, capped raises
        stack = int(round(bb * np.exp(rng.uniform(np.log(0.5), np.log(40)))))
        n_actions = int(rng.integers(3, 7))
        max_raises = int(rng.integers(1, 3))
    else:
        stack = int(min(32767, round(bb * np.exp(rng.uniform(np.log(0.5), np.log(32767 / bb))))))
        n_actions = int(rng.integers(3, 17))
        max_raises = int(rng.integers(0, 5))
    menu = random_menu(rng, n_actions)
    need = max(1, (len(menu) - 1).bit_length())
    hist = int(rng.integers(need, 9))
    cfg = mccfr.menu_cfg(N, stack, menu, max_raises=max_raises, table_log2=int(rng.integers(12, 17)),
                         seed=int(rng.integers(0, 2**63)), small_blind=sb, big_blind=bb,
                         use_cfr_plus=bool(rng.random() < 0.7), use_linear_cfr=bool(rng.random() < 0.7),
                         history_bits=hist)
    if external:
        cfg.sampling = mccfr.SAMPLE_EXTERNAL
    root = None
    if rng.random() < 0.5:
        street = int(rng.integers(1, 4))
        n_active = int(rng.integers(2, N + 1))
        active = sorted(rng.choice(N, n_active, replace=False).tolist())
        cap = stack if external else 32767
        stacks, invested = [], []
        for p in range(N):
            inv = int(rng.integers(0, cap // 2 + 1))
            st = int(rng.integers(0, cap - inv + 1)) if p in active else 0
            if p in active and rng.random() < 0.15:
                st = 0  # already all in
            stacks.append(st)
            invested.append(inv)
        if sum(1 for p in active if stacks[p] > 0) < 1:
            stacks[active[0]] = max(1, stacks[active[0]])
        board = rng.choice(52, street + 2, replace=False).tolist()
        root = mccfr.root_state(street, int(rng.integers(0, N)), active, stacks, invested, board,
                                history=int(rng.integers(0, 1 << 16)))
    start = "random" if rng.random() < 0.6 else "zero"
    return cfg, root, start


def describe(cfg, root, start):
    menu = [(cfg.menu[a].kind, cfg.menu[a].num, cfg.menu[a].den) for a in range(cfg.num_actions)]
    r = "preflop" if root is None else (f"street {root.street} dealer {root.dealer} active {root.active_mask:#x} "
                                       f"stacks {list(root.stack)} invested {list(root.invested)} "
                                       f"board {list(root.board)}")
    return (f"N={cfg.num_players} blinds {cfg.small_blind}/{cfg.big_blind} stack {cfg.starting_stack} "
            f"cap {cfg.max_raises} hist {cfg.history_bits} cfr+ {cfg.use_cfr_plus} lcfr {cfg.use_linear_cfr} "
            f"log2 {cfg.table_log2} sampling {cfg.sampling} {start} menu {menu} root: {r}")


def check_game(hand_table, lut, cfg, root, start, n, lanes, seats_cap):
    row = 8 if cfg.num_actions <= 8 else 16
    with mccfr.Context(mccfr.cfg_3p_20bb(table_log2=4), lanes) as ctx:
        ctx.set_hand_table(hand_table)
        t = ctx.tenant_create(cfg, root)
        if start == "random":
            ctx.tenant_randomize(t, cfg.table_log2)
        seats = list(range(cfg.num_players)) if root is None else \
            [p for p in range(cfg.num_players) if root.active_mask >> p & 1]
        seats = seats[:seats_cap]
        tid = 0
        for L, up in enumerate(seats * 2):
            w = L // len(seats) + 1 if cfg.use_linear_cfr else 1
            R0, S0, sig0 = ctx.tenant_download(t)
            print(f"  launch {L}: up={up} t={w} tid={tid} n={n}", flush=True)
            tr = ctx.run_segments([(t, n, up, w, tid)], trace=True)
            dR, dS, aR, aS, otr = oracle.replay_launch(cfg, hand_table, lut, R0, sig0, up, w, tid, n, root=root)
            for f in FIELDS:
                assert np.array_equal(tr[f], otr[f]), (L, f)
            assert np.array_equal(tr["utility"].view(np.uint32), otr["utility"].view(np.uint32)), L
            R1, S1, sig1 = ctx.tenant_download(t)
            eR, eS = helpers.expected_post(R0, S0, dR, dS, bool(cfg.use_cfr_plus))
            okR, okS = helpers.within_tol(R1, eR, aR), helpers.within_tol(S1, eS, aS)
            assert okR.all(), (L, np.argwhere(~okR)[:5])
            assert okS.all(), (L, np.argwhere(~okS)[:5])
            assert np.array_equal(sig1, oracle.regret_matching(R1, row)), L
            tid += n


@pytest.mark.parametrize("case", range(96))
def test_random_game_outcome_sampling(hand_table, golden_lut, case):
    cfg, root, start = random_game(1000 + case)
    print(f"case {case}: {describe(cfg, root, start)}", flush=True)
    check_game(hand_table, golden_lut, cfg, root, start, n=1024, lanes=1024, seats_cap=4)


@pytest.mark.parametrize("case", range(32))
def test_random_game_external_sampling(hand_table, golden_lut, case):
    cfg, root, start = random_game(2000 + case, external=True)
    print(f"case {case}: {describe(cfg, root, start)}", flush=True)
    check_game(hand_table, golden_lut, cfg, root, start, n=512, lanes=512 * 128, seats_cap=3)


@pytest.mark.parametrize("row", [8, 16])
def test_random_games_batched_equal_each_alone(hand_table, row):
    """Random games of one row width (different table sizes, seat counts,
    roots and update seats) in one launch: every tenant's traces and tables
    bit-identical to the same launch of that tenant alone."""
    games = []
    seed = 3000 + row
    while len(games) < 8:
        cfg, root, start = random_game(seed)
        seed += 1
        if (8 if cfg.num_actions <= 8 else 16) == row:
            games.append((cfg, root, start))
    n = 768
    with mccfr.Context(mccfr.cfg_3p_20bb(table_log2=4), n * len(games)) as ctx:
        ctx.set_hand_table(hand_table)
        ids = [ctx.tenant_create(cfg, root) for cfg, root, _ in games]
        for t, (cfg, root, start) in zip(ids, games):
            print(f"tenant {t}: {describe(cfg, root, start)}", flush=True)
            ctx.tenant_randomize(t, 77 + t)
        pre = [ctx.tenant_download(t) for t in ids]
        for t, (R0, S0, _) in zip(ids, pre):  # both runs start from an upload (a CFR+ upload clamps after the launch)
            ctx.tenant_upload(t, R0, S0)
        segs = []
        for t, (cfg, root, _) in zip(ids, games):
            seats = list(range(cfg.num_players)) if root is None else \
                [p for p in range(cfg.num_players) if root.active_mask >> p & 1]
            segs.append((t, n, seats[t % len(seats)], 2, 500 * t))
        mixed_tr = ctx.run_segments(segs, trace=True)
        mixed = [ctx.tenant_download(t) for t in ids]
        for k, (t, (R0, S0, _)) in enumerate(zip(ids, pre)):
            ctx.tenant_upload(t, R0, S0)
            alone_tr = ctx.run_segments([segs[k]], trace=True)
            R1, S1, sig1 = ctx.tenant_download(t)
            part = mixed_tr[k * n:(k + 1) * n]
            for f in FIELDS + ("utility",):
                assert np.array_equal(part[f].view(np.uint32) if f == "utility" else part[f],
                                      alone_tr[f].view(np.uint32) if f == "utility" else alone_tr[f]), (k, f)
            assert np.array_equal(mixed[k][0], R1), k
            assert np.array_equal(mixed[k][1], S1), k
            assert np.array_equal(mixed[k][2], sig1), k
