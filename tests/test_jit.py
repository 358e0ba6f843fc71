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


def batch_module(db, progs, tab, compile_it=True):
    """vg_jitb_check over progs: (distinct programs, fn per program)."""
    n = len(progs)
    arrs = [_ops(db, p) for p in progs]
    OpP = C.POINTER(va.Op)
    ops = (OpP * n)(*[C.cast(a, OpP) for a, _ in arrs])
    nums = (C.c_uint * n)(*[k for _, k in arrs])
    pcs = (C.c_uint * n)(*[_pc(a, k) for a, k in arrs])
    tabs = (C.POINTER(va.TabletMeta) * n)(*[tab] * n)
    al = (C.c_int * n)(*[1] * n)
    fn = (C.c_uint32 * n)()
    ms = C.c_double(0)
    err = C.create_string_buffer(4096)
    k = va.lib.vg_jitb_check(ops, nums, pcs, tabs, al, n, fn, 1 if compile_it else 0, C.byref(ms), err, 4096)
    assert k > 0, err.value.decode()
    return k, list(fn)


def test_batch_module_dedups_and_compiles(db):
    """A batch module holds each distinct program once: the selectivity
    sweep shares one function, config 1 and the query suite add theirs."""
    tid = db.gen_table("test", 5000, 4096, seed=1)
    tab = db.tablets(tid)[0]
    progs = [config2_sql(t) for t in (0.1, 10.0, 50.0, 90.0)] + [CONFIG1_SQL] + QUERY_SUITE[:3]
    k, fn = batch_module(db, progs, tab)
    assert fn[:4] == [0, 0, 0, 0]
    assert fn[4] == 1
    assert k == len(set(fn)) and k >= 3
