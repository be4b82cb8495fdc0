"""mccfr_run_launch_noalloc (the launch on the library's static scratch)
against mccfr_run_launch (the context's scratch): identical tables after two
calls in a row, the second smaller, from the same seeded table state."""
import numpy as np
import pytest

import mccfr

pytestmark = pytest.mark.gpu


def test_noalloc_launch_equals_run_launch(hand_table):
    cfg = mccfr.cfg_6p_100bb(table_log2=16, seed=21)
    # the allocation-free launch does not use the context's scratch: its
    # context holds scratch for 256 traversals only
    with mccfr.Context(cfg, 16384) as a, mccfr.Context(cfg, 256) as b:
        for c in (a, b):
            c.set_hand_table(hand_table)
            c.randomize_tables(5)
        calls = [(0, 3, 0, 12000), (4, 3, 12000, 3000), (1, 4, 15000, 16384)]
        for up, t, tid, n in calls:
            print(f"noalloc: update player {up} tid {tid} n {n}", flush=True)
            a.run_launch(up, t, tid, n)
            b.run_launch_noalloc(up, t, tid, n)
            Ra, Sa, siga = a.download()
            Rb, Sb, sigb = b.download()
            assert np.array_equal(Ra.view(np.uint32), Rb.view(np.uint32))
            assert np.array_equal(Sa.view(np.uint32), Sb.view(np.uint32))
            assert np.array_equal(siga, sigb)
            assert (Sa != 0).sum() > 1000  # the launches did touch the tables
        rc = mccfr.lib().mccfr_run_launch_noalloc(b.h, 0, 1, 0, 16385)
        assert rc == mccfr.E_ARG
        assert b"16384" in mccfr.lib().mccfr_last_error(b.h)
        assert mccfr.lib().mccfr_run_launch_noalloc(b.h, 0, 1, 0, 0) == mccfr.E_ARG
