"""The specialised kernels' generated code, run on the CPU (tests/kernel_cpu.py)
against the oracle, bit for bit and in row order: the three ways of feeding
the filter (columns staged in LDS; loaded into registers tile by tile and
the ones the result phase needs loaded again for the selected rows; a
one-column filter held in registers for both phases, the default shape's
choice for such filters), over
the benchmark queries, the reference's KATs and suite, typed and cast
programs, ragged layouts, random programs and a capacity that does not
fit.  The same source is what hiprtc compiles for the GPU; the kernel
entry, the chunk counter and the epoch are left to the GPU tests."""
from __future__ import annotations

import ctypes as C
import subprocess

import numpy as np
import pytest

import virginian_amd as va
from oracle import oracle
from tests import kernel_cpu as K
from tests.helpers import (CONFIG1_SQL, KAT_CONDITIONS, QUERY_SUITE, assert_same, cast_programs, config2_sql,
                           simpledb, typed_program, typed_table)
from virginian_amd import workload as W

MODES = ["1", "0", "held"]   # VG_JIT_STAGE: staged in LDS, registers; the default shape

# A random program (seed 5, number 187 of the register-mode sweep) whose
# filter drops the rows failing two OR-ed compares: the compiler once folded
# that mask logic into a v_bitop3 with a wrong truth table (0 rows instead
# of 509 on 4096 rows); kept as a regression case everywhere.
P187 = [(0, 0, 0, 0, None), (1, 2, 0, 0, 'r0'), (2, 0, 0, 27, None), (4, 0, 2, 0, None), (4, 1, 0, 0, None),
        (19, 2, 0, 1, None), (10, 3, 84, 0, None), (12, 2, 3, 14, 1), (4, 4, 2, 0, None), (4, 5, 1, 0, None),
        (16, 4, 5, 14, 1), (4, 6, 0, 0, None), (4, 7, 0, 0, None), (14, 6, 7, 25, 0), (4, 8, 3, 0, None),
        (11, 9, 0, 0, 62.018009185791016), (13, 8, 9, 20, 1), (4, 10, 4, 0, None),
        (11, 11, 0, 0, 39.76091003417969), (14, 10, 11, 25, 0), (4, 12, 2, 0, None), (10, 13, -24, 0, None),
        (12, 12, 13, 24, 1), (8, 0, 0, 0, None), (4, 15, 4, 0, None), (6, 15, 1, 0, None), (7, 0, 0, 0, None),
        (3, 0, 0, 0, None)]


@pytest.fixture(params=MODES, ids=["staged", "registers", "default"])
def mode(request, monkeypatch):
    if request.param == "held":
        monkeypatch.delenv("VG_JIT_STAGE", raising=False)
    else:
        monkeypatch.setenv("VG_JIT_STAGE", request.param)
    return request.param


def _check(db, prog, what="", mode=None):
    want, ill = oracle.select(db, prog)
    assert ill == 0
    got, total, flags, shape = K.select(db, prog)
    assert flags == 0, (what, flags)
    assert total == len(next(iter(want.values()))), (what, total)
    assert_same(got, want, f"{what} {shape}")
    if mode is not None and mode != "held":
        assert shape["staged"] == (1 if mode == "1" else 0) or shape["fbytes_row"] == 0, shape
        assert shape["held"] == 0, shape
    return shape


@pytest.mark.parametrize("rows,tablet_rows", [(1, 0), (4095, 4000), (16385, 5000), (70001, 1 << 14),
                                              (300001, 0)])
def test_layouts(db, mode, rows, tablet_rows):
    """Partial tiles, chunks spanning tablets, tablets smaller than a tile,
    the reference's 8 MB tablets."""
    db.gen_table("test", rows, tablet_rows, seed=rows)
    for sql in (CONFIG1_SQL, config2_sql(50.0), "select id from test where uniformi < 101"):
        _check(db, sql, f"{sql} rows={rows} tablet_rows={tablet_rows}", mode)


def test_config2_thresholds(db, mode):
    db.gen_table("test", 1 << 17, 1 << 16, seed=2)
    for t in (0.1, 10.0, 90.0):
        shape = _check(db, config2_sql(t), f"T={t}", mode)
        # the benchmark query filters on one 4-byte column: the default
        # shape holds it in registers
        assert shape["held"] == (1 if mode == "held" else 0), shape


def test_kats_and_suite(db, mode):
    simpledb(db, 10)
    for q in KAT_CONDITIONS:
        _check(db, q, q)


def test_query_suite(db, mode):
    db.gen_table("test", 50_000, 1 << 14, seed=3)
    for q in QUERY_SUITE:
        _check(db, q, q)


def test_typed_and_casts(db, mode):
    """char / double / int64 columns, the char column first so the others
    are not 16-byte aligned (the per-row load variant)."""
    typed_table(db, 5000, 3)
    for ops in [typed_program()] + cast_programs():
        _check(db, ops, "typed")


@pytest.mark.parametrize("tablet_rows", [0, 1 << 14])
def test_random_programs(db, mode, tablet_rows):
    db.gen_table("test", 40_000, tablet_rows, seed=11)
    rng = np.random.default_rng(177 + tablet_rows)
    for i in range(10):
        sel = float(rng.uniform(0.0001, 0.8)) if rng.random() < 0.7 else None
        prog = W.random_program(rng, max_ops=40, selectivity=sel)
        _check(db, prog, f"program {i}: {prog}")


def test_mask_logic_regression(db, mode):
    db.gen_table("test", 4096, 0, seed=11)
    _check(db, P187, "P187")


def test_capacity(db, mode):
    """A result larger than the capacity: its exact size is reported, the
    overflow flagged, and the rows that fit are the result's first rows."""
    db.gen_table("test", 100_000, 1 << 15, seed=4)
    sql = "select id, uniformf from test where uniformi < 50"
    want, _ = oracle.select(db, sql)
    n = len(want["id"])
    got, total, flags, _ = K.select(db, sql, capacity=1000)
    assert total == n and flags & 1
    assert_same(got, {k: v[:1000] for k, v in want.items()}, "first rows")


def test_no_bitop3_in_compiled_code(tmp_path, monkeypatch):
    """hiprtc builds the specialised kernels without v_bitop3 (the program
    above once compiled to a wrong truth table)."""
    monkeypatch.setenv("VG_JIT_DUMP", str(tmp_path))
    monkeypatch.setenv("VG_JIT_STAGE", "0")
    db = va.Virginian()
    try:
        db.gen_table("test", 4096, 0, seed=11)
        arr, n, keep = va.make_ops(P187)
        pc = [i for i in range(n) if arr[i].op == va.OP_Parallel][0] + 1
        tab = db.tablets(db.table_id("test"))[0]
        err = C.create_string_buffer(2048)
        ms = C.c_double(0)
        assert va.lib.vg_jit_compile_check(arr, n, pc, tab, C.byref(ms), err, 2048) == 1, err.value.decode()
    finally:
        db.close()
    cos = list(tmp_path.glob("*.co"))
    assert cos
    dis = subprocess.run(["/opt/rocm/llvm/bin/llvm-objdump", "-d", str(cos[0])], capture_output=True,
                         text=True, check=True).stdout
    assert "v_cmp" in dis and "v_bitop3" not in dis
