"""The specialised kernels' code generator on the CPU: every program shape
the tests and benchmarks use yields source that hiprtc compiles for gfx950
(no device needed), and the source depends on the program's structure,
not its constants."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import virginian_amd as va
from tests.helpers import (CONFIG1_SQL, KAT_CONDITIONS, QUERY_SUITE, cast_programs, config2_sql, simpledb,
                           typed_program, typed_table)
from virginian_amd import workload as W


def _ops(db, prog):
    if isinstance(prog, str):
        vm = db.compile(prog)
        arr, n = va.ops_array(vm)
        db.free_vm(vm)
        return arr, n
    arr, n, keep = va.make_ops(prog)
    _ops.keep.append(keep)
    return arr, n


_ops.keep = []


def _pc(arr, n):
    return [i for i in range(n) if arr[i].op == va.OP_Parallel][0] + 1


def source(db, prog, tab) -> str:
    arr, n = _ops(db, prog)
    buf = C.create_string_buffer(1 << 21)
    err = C.create_string_buffer(1024)
    L = va.lib.vg_jit_source(arr, n, _pc(arr, n), tab, buf, 1 << 21, err, 1024)
    assert L > 0, err.value.decode()
    return buf.value.decode()


def compiles(db, prog, tab) -> float:
    arr, n = _ops(db, prog)
    err = C.create_string_buffer(2048)
    ms = C.c_double(0)
    ok = va.lib.vg_jit_compile_check(arr, n, _pc(arr, n), tab, C.byref(ms), err, 2048)
    assert ok == 1, err.value.decode()
    return ms.value


def test_constants_share_a_kernel(db):
    """The selectivity sweep differs only in a constant: one source."""
    tid = db.gen_table("test", 5000, 4096, seed=1)
    tab = db.tablets(tid)[0]
    srcs = {source(db, config2_sql(t), tab) for t in (0.1, 10.0, 50.0, 90.0)}
    assert len(srcs) == 1
    assert source(db, CONFIG1_SQL, tab) not in srcs


def test_benchmark_shapes_compile(db):
    tid = db.gen_table("test", 5000, 4096, seed=1)
    tab = db.tablets(tid)[0]
    for prog in [CONFIG1_SQL, config2_sql(10.0)] + QUERY_SUITE[:4]:
        compiles(db, prog, tab)


def test_typed_programs_compile(db):
    typed_table(db, 100, 3)
    tab = db.tablets(db.table_id("test"))[0]
    for prog in [typed_program()] + cast_programs():
        compiles(db, prog, tab)


def test_kat_programs_compile(db):
    simpledb(db, 10)
    tab = db.tablets(db.table_id("test"))[0]
    for q in KAT_CONDITIONS[:3]:
        compiles(db, q, tab)


def test_random_programs_compile(db):
    tid = db.gen_table("test", 5000, 4096, seed=1)
    tab = db.tablets(tid)[0]
    rng = np.random.default_rng(5)
    for _ in range(6):
        prog = W.random_program(rng, max_ops=40, selectivity=float(rng.uniform(0.01, 0.8)))
        compiles(db, prog, tab)
