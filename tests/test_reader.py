"""The result reader (virg_reader_*, reference: src/reader/) on the CPU.

Our reader steps through a chain of result tablets built with the host
library, the way virg_vm_gpu leaves them (the first tablet filled to its
capacity, the rest in one tablet made by virg_vm_allocresult_rows); the
reference's own reader (oracle/_ref) steps through the reference VM's
result chain for the same query.  The two must hand out the same row
bytes call by call, including the reference's quirk that the final row
is copied by the call that returns VIRG_FAIL (src/reader/row.c:42-48), so
a `while (virg_reader_row(...) != VIRG_FAIL)` loop (src/test/sql.cc:257)
sees every row but the last."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import virginian_amd as va
from oracle import ref
from tests.test_oracle import gen_columns

needs_ref = pytest.mark.skipif(not ref.available(), reason="reference VM not built (oracle/build_ref.sh)")

SQL = "select id, uniformi from test where uniformi < 95"
ROW_BYTES = 12          # id (int64) + uniformi (int)
COLS = ("uniformi", "normali5", "normali20", "uniformf", "normalf5", "normalf20")


def _bind():
    lib = va.lib
    PT = C.POINTER(va.TabletMeta)
    V = C.POINTER(va.State)
    for name, res, args in (
            ("virg_vm_allocresult", C.c_int, [V, C.POINTER(va.VM), C.POINTER(PT), PT]),
            ("virg_vm_allocresult_rows", C.c_int, [V, C.POINTER(va.VM), C.POINTER(PT), PT, C.c_uint]),
            ("virg_tablet_addcolumn", C.c_int, [PT, C.c_char_p, C.c_int]),
            ("virg_tablet_addmaxrows", C.c_int, [V, PT]),
            ("virg_tablet_unlock", C.c_int, [V, C.c_uint]),
            ("virg_reader_init", C.c_int, [V, C.POINTER(va.Reader), C.POINTER(va.VM)]),
            ("virg_reader_free", C.c_int, [V, C.POINTER(va.Reader)])):
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


def _fill(tab, cols: list[np.ndarray], start: int, n: int) -> None:
    t = tab.contents
    base = C.addressof(t)
    for j, c in enumerate(cols):
        part = np.ascontiguousarray(c[start:start + n])
        C.memmove(base + t.fixed_block + t.fixed_offset[j] + t.rows * t.fixed_stride[j], part.ctypes.data,
                  part.nbytes)
    t.rows += n


def our_chain(db, ids: np.ndarray, vals: np.ndarray):
    """A vm whose result chain holds (ids, vals): the first tablet filled to
    capacity, the rest in one virg_vm_allocresult_rows tablet."""
    lib = _bind()
    vm = lib.virg_vm_init()
    first = C.POINTER(va.TabletMeta)()
    assert lib.virg_vm_allocresult(db.ptr, vm, C.byref(first), None) == va.VIRG_SUCCESS
    assert lib.virg_tablet_addcolumn(first, b"id", va.VIRG_INT64) == va.VIRG_SUCCESS
    assert lib.virg_tablet_addcolumn(first, b"uniformi", va.VIRG_INT) == va.VIRG_SUCCESS
    assert lib.virg_tablet_addmaxrows(db.ptr, first) == va.VIRG_SUCCESS
    cap = first.contents.possible_rows
    assert 0 < cap < len(ids), "the result must span two tablets"
    _fill(first, [ids, vals], 0, cap)
    rest = C.POINTER(va.TabletMeta)()
    assert lib.virg_vm_allocresult_rows(db.ptr, vm, C.byref(rest), first, len(ids) - cap) == va.VIRG_SUCCESS
    _fill(rest, [ids, vals], cap, len(ids) - cap)
    for t in (first, rest):
        lib.virg_tablet_unlock(db.ptr, t.contents.id)
    return vm


def read_all(db, vm) -> tuple[np.ndarray, int]:
    """Our reader: the buffer after every virg_reader_row call up to and
    including the first VIRG_FAIL, and the calls that succeeded."""
    lib = _bind()
    r = va.Reader()
    assert lib.virg_reader_init(db.ptr, C.byref(r), vm) == va.VIRG_SUCCESS
    out, ok = [], 0
    while True:
        rc = lib.virg_reader_row(db.ptr, C.byref(r))
        # (a c_char array field reads as bytes up to its first NUL: take the memory)
        out.append(C.string_at(C.addressof(r) + va.Reader.buffer.offset, ROW_BYTES))
        if rc == va.VIRG_FAIL:
            break
        ok += 1
    lib.virg_reader_free(db.ptr, C.byref(r))
    return np.frombuffer(b"".join(out), dtype=np.uint8).reshape(-1, ROW_BYTES), ok


def rows_bytes(ids: np.ndarray, vals: np.ndarray) -> np.ndarray:
    rec = np.empty(len(ids), dtype=[("id", "<i8"), ("v", "<i4")])
    rec["id"], rec["v"] = ids, vals
    return rec.view(np.uint8).reshape(-1, ROW_BYTES)


@needs_ref
def test_reader_matches_reference_reader(db):
    n = 500_000
    c = gen_columns(n, 3)
    r = ref.RefDB()
    try:
        r.table("test", c["id"], [(k, c[k]) for k in COLS])
        ref_rows, ref_ok = r.run_reader(oracle_ops(db, c, n), ROW_BYTES, n + 1)
    finally:
        r.close()
    sel = c["uniformi"] < 95
    ids, vals = c["id"][sel], c["uniformi"][sel]
    want = rows_bytes(ids, vals)
    # the reference reader: every row, the last one by the failing call
    assert len(ref_rows) == len(ids) and ref_ok == len(ids) - 1
    assert np.array_equal(ref_rows, want)
    vm = our_chain(db, ids, vals)
    try:
        ours, ok = read_all(db, vm)
    finally:
        va.lib.virg_vm_freeresults(db.ptr, vm)
        va.lib.virg_vm_cleanup(db.ptr, vm)
    assert ok == ref_ok
    assert np.array_equal(ours, ref_rows)


def oracle_ops(db, c, n):
    """The op program of SQL, compiled by our front-end (its op programs
    equal the reference generator's: tests/test_sql.py)."""
    db.create_table("test", va.VIRG_INT64, [(k, va.VIRG_INT if k.endswith("i") or "normali" in k
                                             else va.VIRG_FLOAT) for k in COLS])
    vm = db.compile(SQL)
    arr, k = va.ops_array(vm)
    db.free_vm(vm)
    return arr, k


def test_reader_single_tablet_and_getrows(db):
    """One result tablet, a handful of rows: SUCCESS for all but the last
    call, which copies the last row and returns FAIL; getrows counts all."""
    lib = _bind()
    lib.virg_reader_getrows.argtypes = [C.POINTER(va.State), C.POINTER(va.Reader), C.POINTER(C.c_uint)]
    vm = lib.virg_vm_init()
    t = C.POINTER(va.TabletMeta)()
    assert lib.virg_vm_allocresult(db.ptr, vm, C.byref(t), None) == va.VIRG_SUCCESS
    lib.virg_tablet_addcolumn(t, b"id", va.VIRG_INT64)
    lib.virg_tablet_addcolumn(t, b"uniformi", va.VIRG_INT)
    lib.virg_tablet_addmaxrows(db.ptr, t)
    ids = np.arange(5, dtype=np.int64) * 7
    vals = np.arange(5, dtype=np.int32) - 2
    _fill(t, [ids, vals], 0, 5)
    lib.virg_tablet_unlock(db.ptr, t.contents.id)
    r = va.Reader()
    assert lib.virg_reader_init(db.ptr, C.byref(r), vm) == va.VIRG_SUCCESS
    rows = C.c_uint(0)
    assert lib.virg_reader_getrows(db.ptr, C.byref(r), C.byref(rows)) == va.VIRG_SUCCESS
    assert rows.value == 5
    lib.virg_reader_free(db.ptr, C.byref(r))
    got, ok = read_all(db, vm)
    assert ok == 4 and np.array_equal(got, rows_bytes(ids, vals))
    va.lib.virg_vm_freeresults(db.ptr, vm)
    va.lib.virg_vm_cleanup(db.ptr, vm)
