"""The allocation-free entry point (vg_select_noalloc) against the
prepared-query path (vg_query_run) it is a variant of: bit for bit, on two
calls in a row through the C ABI, the second over fewer rows, so the
shared scratch keeps the first call's contents when the second starts."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import virginian_amd as va
from tests.helpers import CONFIG1_SQL, assert_same, config2_sql

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def host_columns(db, table: str):
    """The table's key and fixed columns, concatenated over its tablets."""
    tabs = db.tablets(db.table_id(table))
    t0 = tabs[0].contents
    ncols = t0.fixed_columns
    types = [t0.fixed_type[c] for c in range(ncols)]
    key, cols = [], [[] for _ in range(ncols)]
    for tp in tabs:
        t = tp.contents
        base = C.addressof(t)
        n = t.rows
        kb = (C.c_char * (n * t.key_stride)).from_address(base + t.key_block)
        key.append(np.frombuffer(kb, dtype=va.NP_TYPE[t.key_type]).copy())
        for c in range(ncols):
            cb = (C.c_char * (n * t.fixed_stride[c])).from_address(base + t.fixed_block + t.fixed_offset[c])
            cols[c].append(np.frombuffer(cb, dtype=va.NP_TYPE[types[c]]).copy())
    return np.concatenate(key), [np.concatenate(c) for c in cols], types, t0.key_type


class Device:
    """Device buffers made by the caller (the entry point makes none)."""

    def __init__(self, db):
        self.db, self.ptrs = db, []

    def alloc(self, nbytes: int) -> int:
        p = va.lib.vg_device_alloc(self.db.ptr, max(nbytes, 256))
        assert p
        self.ptrs.append(p)
        return p

    def put(self, a: np.ndarray) -> int:
        p = self.alloc(a.nbytes)
        assert hip().hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def get(self, p: int, n: int, dtype) -> np.ndarray:
        out = np.empty(n, dtype=dtype)
        if n:
            assert hip().hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            va.lib.vg_device_free(self.db.ptr, p)


def noalloc_select(db, dev: Device, sql: str, dcols, types, dkey, key_type, rows: int, capacity: int):
    vm = db.compile(sql)
    arr, n = va.ops_array(vm)
    names, rtypes = va.result_schema(vm)
    db.free_vm(vm)
    outs = [dev.alloc(capacity * np.dtype(va.NP_TYPE[t]).itemsize) for t in rtypes]
    result = dev.alloc(32)
    err = C.create_string_buffer(256)
    rc = va.lib.vg_select_noalloc(arr, n, (C.c_void_p * len(dcols))(*dcols), (C.c_int * len(types))(*types),
                                  len(types), dkey, key_type, rows, (C.c_void_p * len(outs))(*outs), capacity,
                                  result, None, err, 256)
    assert rc == va.VIRG_SUCCESS, err.value.decode()
    assert hip().hipDeviceSynchronize() == 0
    res = dev.get(result, 4, np.uint64)
    assert res[1] & 0xFFFFFFFF == 0, "capacity exceeded"
    nsel = int(res[0])
    return {nm: dev.get(o, nsel, va.NP_TYPE[t]) for nm, o, t in zip(names, outs, rtypes)}


def query_run(db, sql: str) -> dict:
    q = db.prepare(sql, flags=va.VG_QUERY_INTERP)
    try:
        q.run()
        return q.fetch()
    finally:
        q.close()


def test_noalloc_matches_query_run(gpu_db):
    big, small = 1 << 20, 300_001
    gpu_db.gen_table("big", big, 1 << 18, seed=21)
    gpu_db.gen_table("small", small, 1 << 18, seed=21)    # the same rows as big's first ones
    key, cols, types, key_type = host_columns(gpu_db, "big")
    skey, scols, _, _ = host_columns(gpu_db, "small")
    assert np.array_equal(skey, key[:small])
    assert all(np.array_equal(s.view(np.uint8), c[:small].view(np.uint8)) for s, c in zip(scols, cols))
    dev = Device(gpu_db)
    try:
        dkey = dev.put(key)
        dcols = [dev.put(c) for c in cols]
        for sql in (config2_sql(50.0).replace("from test", "from big"), CONFIG1_SQL.replace("from test", "from big")):
            first = noalloc_select(gpu_db, dev, sql, dcols, types, dkey, key_type, big, big)
            want = query_run(gpu_db, sql)
            assert_same(first, want, f"first call: {sql}")
            sql_s = sql.replace("from big", "from small")
            second = noalloc_select(gpu_db, dev, sql_s, dcols, types, dkey, key_type, small, small)
            want_s = query_run(gpu_db, sql_s)
            assert_same(second, want_s, f"second, smaller call: {sql_s}")
            assert 0 < len(second[list(second)[0]]) < len(first[list(first)[0]])
    finally:
        dev.free()


def test_noalloc_refuses_beyond_its_scratch(gpu_db):
    gpu_db.gen_table("t", 1000, 0, seed=1)
    vm = gpu_db.compile("select id from t where uniformi < 50")
    arr, n = va.ops_array(vm)
    gpu_db.free_vm(vm)
    err = C.create_string_buffer(256)
    types = (C.c_int * 6)(*[va.VIRG_INT] * 3, *[va.VIRG_FLOAT] * 3)
    cols = (C.c_void_p * 6)(*[None] * 6)
    rc = va.lib.vg_select_noalloc(arr, n, cols, types, 6, None, va.VIRG_INT64, (16 << 20) + 1,
                                  (C.c_void_p * 1)(None), 0, None, None, err, 256)
    assert rc == va.VIRG_FAIL
    assert b"at most 16777216 rows" in err.value
