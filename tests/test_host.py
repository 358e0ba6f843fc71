"""Host-side library (no GPU): tables, tablets and their layout, the
reference insert path against bulk tablets, argument checking."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import virginian_amd as va
from tests.helpers import simpledb


def test_duplicate_table_rejected(db):
    simpledb(db, 3)
    with pytest.raises(va.VirginianError):
        db.create_table("test", va.VIRG_INT, [("x", va.VIRG_INT)])
    assert db.numrows(db.table_id("test")) == 3


def test_append_tablet_checks_arrays(db):
    tid = db.create_table("t", va.VIRG_INT64, [("a", va.VIRG_INT), ("b", va.VIRG_FLOAT)])
    key = np.arange(10, dtype=np.int64)
    with pytest.raises(va.VirginianError):
        db.append_tablet(tid, key, [np.zeros(10, np.int32)])                     # one column short
    with pytest.raises(va.VirginianError):
        db.append_tablet(tid, key, [np.zeros(10, np.int64), np.zeros(10, np.float32)])  # width
    with pytest.raises(va.VirginianError):
        db.append_tablet(tid, key, [np.zeros(9, np.int32), np.zeros(10, np.float32)])   # length
    db.append_tablet(tid, key, [np.arange(10, dtype=np.int32), np.ones(10, np.float32)])
    assert db.numrows(tid) == 10


@pytest.mark.parametrize("rows,tablet_rows", [(1, 0), (300_000, 0), (70_001, 1000), (5, 1 << 20)])
def test_tablet_layout(db, rows, tablet_rows):
    """Every tablet passes the reference's layout check (src/tablet/check.c:
    17-57), rows add up, the key is the row index, and the reference's
    8 MB tablets (tablet_rows = 0, the row-by-row insert path) are chained
    in order."""
    db.gen_table("test", rows, tablet_rows, seed=2)
    tid = db.table_id("test")
    tabs = db.tablets(tid)
    assert sum(t.contents.rows for t in tabs) == rows == db.numrows(tid)
    row0 = 0
    for k, t in enumerate(tabs):
        m = t.contents
        assert va.lib.virg_tablet_check(t) == va.VIRG_SUCCESS
        assert m.last_tablet == (k == len(tabs) - 1)
        if tablet_rows == 0:
            assert m.size <= va.VIRG_TABLET_SIZE
        base = C.cast(t, C.c_void_p).value
        key = np.ctypeslib.as_array((C.c_int64 * m.rows).from_address(base + m.key_block))
        assert np.array_equal(key, np.arange(row0, row0 + m.rows))
        row0 += m.rows


def test_insert_path_matches_bulk(db):
    """Row-by-row virg_table_insert (the reference's ingest path) and bulk
    tablets hold the same values."""
    a, b = va.Virginian(), va.Virginian()
    try:
        a.gen_table("test", 20_000, 0, seed=9)
        b.gen_table("test", 20_000, 1 << 20, seed=9)

        def column(d, j):
            out = []
            for t in d.tablets(d.table_id("test")):
                m = t.contents
                base = C.cast(t, C.c_void_p).value + m.fixed_block + m.fixed_offset[j]
                out.append(np.ctypeslib.as_array((C.c_uint8 * (m.rows * m.fixed_stride[j])).from_address(base)).copy())
            return np.concatenate(out)

        for j in range(6):
            assert np.array_equal(column(a, j), column(b, j)), j
    finally:
        a.close()
        b.close()


def test_sizes_and_generalize():
    """virg_sizeof and virg_generalizetype (src/util.c:14-26: equal types
    stay, numeric types give the larger type number; char mixed with
    another type is an assert in the reference, a null type here)."""
    assert [va.lib.virg_sizeof(t) for t in range(5)] == [4, 8, 4, 8, 1]
    va.lib.virg_generalizetype.restype = C.c_int
    for x in range(5):
        for y in range(5):
            g = va.lib.virg_generalizetype(x, y)
            if x == y:
                assert g == x
            elif x <= va.VIRG_DOUBLE and y <= va.VIRG_DOUBLE:
                assert g == max(x, y)
            else:
                assert g > va.VIRG_CHAR
