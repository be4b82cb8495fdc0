"""Host-side hand-table checks (mccfr_validate_hand_table[_file]): the
reference's size and royal-flush checks (hand_eval.cpp:59-86,
cfr_gpu.cu:270-299) plus a range check on every word, all before any device
work.  No GPU: these entry points are host code."""
import os
import subprocess

import numpy as np
import pytest

import mccfr

WORDS = mccfr.HAND_TABLE_WORDS
ROYAL = [51, 50, 49, 48, 47, 0, 14]  # As Ks Qs Js Ts 2c 3d


def rm(c):
    return (c % 13) * 4 + c // 13 + 1


def royal_word_index(hr):
    """Index of the word the royal flush's seventh lookup reads."""
    p = int(hr[53 + rm(ROYAL[0])])
    for c in ROYAL[1:6]:
        p = int(hr[p + rm(c)])
    return p + rm(ROYAL[6])


@pytest.fixture(scope="module")
def table(hand_table):
    return np.array(hand_table, copy=True)


def test_generated_table_is_accepted(table):
    assert mccfr.validate_hand_table(table) == 0


def test_wrong_size_is_rejected(table):
    assert mccfr.validate_hand_table(table[:-1]) == mccfr.E_TABLE
    assert b"32,487,834" in mccfr.lib().mccfr_last_error(None)


def test_bad_royal_flush_is_rejected(table):
    i = royal_word_index(table)
    saved = int(table[i])
    assert (saved >> 12) == 9  # straight flush category, ordinal 10
    try:
        table[i] = (9 << 12) | 9  # a lower straight flush
        assert mccfr.validate_hand_table(table) == mccfr.E_TABLE
        assert b"7462" in mccfr.lib().mccfr_last_error(None)
    finally:
        table[i] = saved
    assert mccfr.validate_hand_table(table) == 0


@pytest.mark.parametrize("bad", [WORDS - 52, WORDS, 1 << 30, -1, -(1 << 31)])
def test_out_of_range_word_is_rejected(table, bad):
    """A word w that a walk could use as a state base reads hr[w + c] for c up
    to 52; any w outside [0, words - 53] is rejected, wherever it sits (here in
    the unused header, the first state block and the last word)."""
    for i in (0, 60, WORDS - 1):
        saved = int(table[i])
        try:
            table[i] = bad
            assert mccfr.validate_hand_table(table) == mccfr.E_TABLE
            assert b"outside" in mccfr.lib().mccfr_last_error(None)
        finally:
            table[i] = saved
    table[7] = WORDS - 53  # the largest legal word
    try:
        assert mccfr.validate_hand_table(table) == 0
    finally:
        table[7] = 0


def test_file_checks(tmp_path, table):
    good = tmp_path / "good.dat"
    table.tofile(good)
    assert mccfr.validate_hand_table_file(str(good)) == 0
    short = tmp_path / "truncated.dat"
    table[: WORDS - 1000].tofile(short)
    assert mccfr.validate_hand_table_file(str(short)) == mccfr.E_TABLE
    assert b"129,951,336 bytes" in mccfr.lib().mccfr_last_error(None)
    empty = tmp_path / "empty.dat"
    empty.write_bytes(b"")
    assert mccfr.validate_hand_table_file(str(empty)) == mccfr.E_TABLE
    corrupt = tmp_path / "corrupt.dat"
    t2 = table.copy()
    t2[royal_word_index(t2) - 5] = WORDS + 7
    t2.tofile(corrupt)
    assert mccfr.validate_hand_table_file(str(corrupt)) == mccfr.E_TABLE
    assert mccfr.validate_hand_table_file(str(tmp_path / "missing.dat")) == mccfr.E_IO
    assert mccfr.lib().mccfr_validate_hand_table_file(None) == mccfr.E_ARG


GEN_TABLE = os.path.join(os.path.dirname(mccfr.LIB_PATH), "gen_table")


def run_gen_table(path):
    if not os.path.exists(GEN_TABLE):
        pytest.fail(f"{GEN_TABLE} is missing: build with `make -C {os.path.dirname(GEN_TABLE)}`")
    return subprocess.run([GEN_TABLE, str(path)], capture_output=True, text=True, timeout=300)


def test_gen_table_writes_the_generated_table(tmp_path, table):
    """The build-time generator (`gen_table [output_path]`, the reference's
    tools/gen_lookup_table.cpp:63-66 invocation) writes the same words as
    the in-memory generator; a valid file is kept; a file that fails the
    load checks is written again."""
    out = tmp_path / "data" / "handranks.dat"
    out.parent.mkdir()
    r = run_gen_table(out)
    assert r.returncode == 0, r.stderr
    assert "wrote" in r.stdout
    words = np.fromfile(out, dtype="<i4")
    assert words.size == WORDS and np.array_equal(words, table)
    assert mccfr.validate_hand_table_file(str(out)) == 0
    r = run_gen_table(out)
    assert r.returncode == 0 and "present and valid" in r.stdout
    words[royal_word_index(words) - 5] = WORDS + 7  # a state word outside the table
    words.tofile(out)
    assert mccfr.validate_hand_table_file(str(out)) == mccfr.E_TABLE
    r = run_gen_table(out)
    assert r.returncode == 0 and "wrote" in r.stdout, r.stderr
    assert np.array_equal(np.fromfile(out, dtype="<i4"), table)
    assert not os.path.exists(str(out) + ".tmp")


def test_gen_table_reports_an_unwritable_path(tmp_path):
    r = run_gen_table(tmp_path / "no_such_dir" / "handranks.dat")
    assert r.returncode == 1 and "cannot create" in r.stderr
