"""mccfr_load_hand_table / mccfr_set_hand_table refuse bad tables with
MCCFR_E_TABLE before any device work (hand_eval.cpp:59-86,
cfr_gpu.cu:1086-1104), and the context stays usable afterwards."""
import numpy as np
import pytest

import mccfr

pytestmark = pytest.mark.gpu


def test_bad_tables_are_refused_and_the_context_stays_usable(tmp_path, hand_table):
    short = tmp_path / "short.dat"
    hand_table[:1000].tofile(short)
    corrupt = np.array(hand_table, copy=True)
    corrupt[100] = mccfr.HAND_TABLE_WORDS  # a state word past the table's end
    with mccfr.Context(mccfr.reference_cfg(num_players=3, starting_stack=400, table_log2=12, seed=3), 1024) as ctx:
        rc = mccfr.lib().mccfr_load_hand_table(ctx.h, str(short).encode())
        assert rc == mccfr.E_TABLE
        assert b"129,951,336" in mccfr.lib().mccfr_last_error(ctx.h)
        with pytest.raises(mccfr.MccfrError) as e:
            ctx.set_hand_table(corrupt)
        assert e.value.code == mccfr.E_TABLE
        assert mccfr.lib().mccfr_load_hand_table(ctx.h, str(tmp_path / "none.dat").encode()) == mccfr.E_IO
        # no table yet: a launch is refused, not run
        assert mccfr.lib().mccfr_run_launch(ctx.h, 0, 1, 0, 256, None) == mccfr.E_STATE
        good = tmp_path / "good.dat"
        hand_table.tofile(good)
        ctx.load_hand_table(str(good))
        # royal flush; four deuces with a three (the lowest quads: 7296 + 1)
        cards = np.array([[51, 50, 49, 48, 47, 0, 14], [0, 13, 26, 39, 1, 14, 27]], np.uint8)
        assert ctx.eval7(cards, 0).tolist() == [7462, 7297]
        assert ctx.eval7(cards, 1).tolist() == [7462, 7297]
        tr = ctx.run_launch(0, 1, 0, 256, trace=True)
        assert (tr["terminal"] > 0).all()
