"""The large-table configuration at its stated size: 6 players, 200bb (4,000
chips), the 10-action menu {F, C/C, 1/4, 1/2, 3/4, 1, 3/2, 2, 3 pot, all-in},
at most 4 raises per street, 2^28-slot tables with 16-wide rows (2^32 regret
and strategy-sum entries, 32 GiB FP32 + 8 GiB BF16 snapshot) and 2^22
traversals per launch.  Row offsets reach 2^28 * 128 B = 32 GiB, far past the
4 GiB that 32-bit `action * table_size + idx` indexing (cfr_gpu.cu:333, :518,
:788) can address.  One launch from a randomized regret state: every lane's
record bit-exact against the oracle, every touched table entry within
tolerance of the oracle's FP64 sums, every other entry unchanged."""
import numpy as np
import pytest

import helpers
import mccfr
import oracle

pytestmark = pytest.mark.gpu


def download(ctx, regrets=True, sums=True, sigma=True):
    n = ctx.slots * ctx.row
    R = np.empty(n, np.float32) if regrets else None
    S = np.empty(n, np.float32) if sums else None
    sig = np.empty(n, np.uint16) if sigma else None
    ptr = lambda a: None if a is None else a.ctypes.data
    ctx.check(mccfr.lib().mccfr_download_tables(ctx.h, ptr(R), ptr(S), ptr(sig), n, 0))
    return R, S, sig


@pytest.mark.timeout(900)
def test_large_table_2pow28_slots_2pow22_lanes(hand_table, golden_lut):
    cfg = mccfr.cfg_6p_200bb_10(table_log2=28, seed=9)
    n = 1 << 22
    up, tw, tid = 3, 5, 7 << 22
    with mccfr.Context(cfg, n) as ctx:
        assert (ctx.slots, ctx.row) == (1 << 28, 16)
        ctx.set_hand_table(hand_table)
        ctx.randomize_tables(3)
        print("large table: downloading the pre-launch regrets and snapshot", flush=True)
        R0, _, sig0 = download(ctx, sums=False)
        print("large table: one launch of 2^22 traversals", flush=True)
        tr = ctx.run_launch(up, tw, tid, n, trace=True)
        print("large table: oracle replay", flush=True)
        idx, dR, dS, aR, aS, otr = oracle.replay_sparse(cfg, hand_table, golden_lut, R0, sig0, up, tw, tid, n)
        del sig0
        for f in ("path_hash", "decisions", "updates", "terminal", "update_legal"):
            assert np.array_equal(tr[f], otr[f]), f
        assert np.array_equal(tr["utility"].view(np.uint32), otr["utility"].view(np.uint32))
        assert (idx // 16).max() >= 1 << 25  # rows beyond byte offset 4 GiB were updated
        print("large table: downloading the updated tables", flush=True)
        R1, S1, sig1 = download(ctx)
        slots = np.unique(idx // 16)
        rows = R1.reshape(-1, 16)[slots]
        assert np.array_equal(sig1.reshape(-1, 16)[slots].reshape(-1), oracle.regret_matching(rows.reshape(-1), 16))
        del sig1, rows
        S0 = np.zeros(1, np.float32)  # randomized tables start with zero strategy sums
        S0 = np.broadcast_to(S0, R0.shape)
        helpers.check_launch_sparse(R0, S0, R1, S1, idx, dR, dS, aR, aS)
