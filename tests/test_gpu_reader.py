"""The reader over GPU results: virg_query (SQL -> virg_vm_execute ->
virg_vm_gpu, results in host result tablets) then virg_reader_row, row by
row against the oracle, over config 1 and over results spanning several
result tablets; the reference's last-row behaviour (src/reader/row.c:42-48,
src/test/sql.cc:257-265): the final row is copied by the call that returns
VIRG_FAIL, and virg_reader_getrows counts every row."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import virginian_amd as va
from oracle import oracle
from tests.helpers import CONFIG1_SQL

pytestmark = pytest.mark.gpu


def _reader_rows(db, sql: str, dtype):
    lib = va.lib
    lib.virg_reader_getrows.argtypes = [C.POINTER(va.State), C.POINTER(va.Reader), C.POINTER(C.c_uint)]
    rp = C.POINTER(va.Reader)()
    assert lib.virg_query(db.ptr, C.byref(rp), sql.encode()) == va.VIRG_SUCCESS
    n = C.c_uint(0)
    assert lib.virg_reader_getrows(db.ptr, rp, C.byref(n)) == va.VIRG_SUCCESS
    tablets = 0
    node = rp.contents.vm.contents.head_result
    while node:
        tablets += 1
        node = node.contents.next
    size = np.dtype(dtype).itemsize
    buf = bytearray()
    ok = 0
    while True:
        rc = lib.virg_reader_row(db.ptr, rp)
        buf += C.string_at(C.addressof(rp.contents) + va.Reader.buffer.offset, size)
        if rc == va.VIRG_FAIL:
            break
        ok += 1
    lib.virg_release(db.ptr, rp)
    return np.frombuffer(bytes(buf), dtype=dtype), ok, n.value, tablets


def _check(db, sql, dtype, names):
    rows, ok, counted, tablets = _reader_rows(db, sql, dtype)
    want, ill = oracle.select(db, sql)
    assert ill == 0
    m = len(want[names[0]])
    assert counted == m
    assert ok == m - 1 and len(rows) == m, "every row, the last copied by the failing call"
    for name, field in zip(names, dtype.names):
        assert np.array_equal(rows[field].view(f"u{rows[field].itemsize}"),
                              want[name].view(f"u{want[name].itemsize}")), name
    return tablets


def test_reader_config1(gpu_db):
    gpu_db.gen_table("test", 1 << 20, 0, seed=1)
    dt = np.dtype([("id", "<i8"), ("u", "<i4"), ("f", "<f4")])
    _check(gpu_db, CONFIG1_SQL, dt, ["id", "uniformi", "normalf5"])


def test_reader_many_result_tablets(gpu_db, tmp_path):
    """With a database file open the results go to 8 MB result tablets:
    a 950k-row result takes several."""
    gpu_db.create_db(tmp_path / "r.db")
    gpu_db.gen_table("test", 1 << 20, 0, seed=2)
    dt = np.dtype([("id", "<i8"), ("u", "<i4"), ("f", "<f4")])
    tablets = _check(gpu_db, "select id, uniformi, uniformf from test where uniformi < 91", dt,
                     ["id", "uniformi", "uniformf"])
    assert tablets >= 3


def test_reader_in_memory_chain(gpu_db):
    """Without a file the first result tablet is 8 MB and the rest of a
    large result goes into one tablet sized for it."""
    gpu_db.gen_table("test", 1 << 20, 1 << 18, seed=3)
    dt = np.dtype([("id", "<i8"), ("f", "<f4")])
    tablets = _check(gpu_db, "select id, normalf20 from test where normalf20 < 30.0", dt, ["id", "normalf20"])
    assert tablets == 2
