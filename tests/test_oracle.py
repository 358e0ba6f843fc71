"""The CPU oracle pinned on its own, with no GPU.

What fixes the oracle's answers here is not the oracle's code:
  * the reference's known answers (src/test/sql.cc: condition row counts
    and members, folded expressions, column order);
  * numpy brute force: every published query written directly as numpy
    predicates and projections over the generated columns (int32 wrap,
    IEEE float32 round-to-nearest per operation, right-folded AND/OR as
    the reference's grammar builds it, rows in ascending row order);
  * invariants: the multi-threaded ordered run equals the single-threaded
    one, the unordered run (the reference's multi-core VM) is the same
    multiset, results are independent of the tablet size.
The oracle's result for a program must match all of them bit for bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import virginian_amd as va
from oracle import oracle
from tests.helpers import (CONFIG1_SQL, KAT_CONDITION_ROWS, KAT_CONDITIONS, QUERY_SUITE,
                           assert_same, config2_sql, simpledb)

ROWS = 150_000


def gen_columns(rows: int, seed: int) -> dict:
    """The generated table as numpy arrays, straight from the generator."""
    cols = {"id": np.empty(rows, np.int64), "uniformi": np.empty(rows, np.int32),
            "normali5": np.empty(rows, np.int32), "normali20": np.empty(rows, np.int32),
            "uniformf": np.empty(rows, np.float32), "normalf5": np.empty(rows, np.float32),
            "normalf20": np.empty(rows, np.float32)}
    va.lib.virg_gen_rows_mt(seed, 0, rows, *[c.ctypes.data for c in cols.values()], 4)
    return cols


def f32(x):
    return np.float32(x)


def brute_force(c: dict) -> list[tuple[str, np.ndarray, list[str]]]:
    """(sql, selected-row mask, projected column names) for the published
    suite, written directly in numpy."""
    ui, n5, n20 = c["uniformi"], c["normali5"], c["normali20"]
    uf, f5, f20 = c["uniformf"], c["normalf5"], c["normalf20"]
    with np.errstate(over="ignore"):
        ip = n5 * n20                      # int32 products, wrapping
    fp = f5 * f20                          # float32 products
    return [
        (QUERY_SUITE[0], (ui > 60) & (n5 < 0), ["id", "uniformi", "normali5"]),
        (QUERY_SUITE[1], (uf > f32(60.0)) & (f5 < f32(0.0)), ["id", "uniformf", "normalf5"]),
        (QUERY_SUITE[2], (ui > -60) & (n5 < 5), ["id", "uniformi", "normali5"]),
        (QUERY_SUITE[3], (uf > f32(-60.0)) & (f5 < f32(5.0)), ["id", "uniformf", "normalf5"]),
        (QUERY_SUITE[4], (n20 + np.int32(40)) > (ui - np.int32(10)), ["id", "uniformi", "normali20"]),
        (QUERY_SUITE[5], (f20 + f32(40.0)) > (uf - f32(10.0)), ["id", "uniformf", "normalf20"]),
        (QUERY_SUITE[6], (ip >= -5) & (ip <= 5), ["id", "normali5", "normali20"]),
        (QUERY_SUITE[7], (fp >= f32(-5.0)) & (fp <= f32(5.0)), ["id", "normalf5", "normalf20"]),
        # A and B or C and D or E and F  ==  A and (B or (C and (D or (E and F))))
        (QUERY_SUITE[8], (ui >= -1) & ((ui <= 1) | ((n5 >= -1) & ((n5 <= 1) | ((n20 >= -1) & (n20 <= 1))))),
         ["id", "uniformi", "normali5", "normali20"]),
        (QUERY_SUITE[9], (uf >= f32(-1.0)) & ((uf <= f32(1.0)) | ((f5 >= f32(-1.0)) & (
            (f5 <= f32(1.0)) | ((f20 >= f32(-1.0)) & (f20 <= f32(1.0)))))),
         ["id", "uniformf", "normalf5", "normalf20"]),
        (CONFIG1_SQL, (ui < 50) & (n5 > 0), ["id", "uniformi", "normalf5"]),
    ]


@pytest.fixture(scope="module")
def table():
    db = va.Virginian()
    db.gen_table("test", ROWS, 1 << 16, seed=7)
    yield db, gen_columns(ROWS, 7)
    db.close()


# ------------------------------------------------------------ known answers
def test_kat_conditions(db):
    """reference: src/test/sql.cc:241-274 (ConditionsSingle)."""
    simpledb(db, 10)
    for i, q in enumerate(KAT_CONDITIONS):
        got, ill = oracle.select(db, q)
        assert ill == 0
        assert len(got["col0"]) == KAT_CONDITION_ROWS[i], q
        vals = got["col0"].tolist()
        assert vals == sorted(vals)
        if i == 0:
            assert set(vals) <= {3, 7, 8}
        if i == 1:
            assert set(vals) <= {8, 9}


def test_kat_expressions(db):
    """reference: src/test/sql.cc:157-177 (=30), :179-200 (=100), :394-405 ({0,1})."""
    simpledb(db, 1)
    got, _ = oracle.select(db, "select col0 + 10 * (1 + 2) from test")
    assert list(got.values())[0].tolist() == [30]
    got, _ = oracle.select(db, "select col0 + 5 * 20 from test")
    assert got["(col0+(5*20))"].tolist() == [100]
    got, _ = oracle.select(db, "select col0, col1 from test")
    assert got["col0"].tolist() == [0] and got["col1"].tolist() == [1]


def test_simpledb_brute_force(db):
    """All nine reference conditions against numpy on a 1000-row simple
    table (col0 = i), right-folded."""
    simpledb(db, 1000)
    x = np.arange(1000, dtype=np.int32)
    masks = [
        (x < 9) & ((x >= 7) | (x == 3)),
        (x >= 8) | ((x == 3) & (x < 1)),
        (x == 8) & (x == 8),
        (x == 8) | (x == 8),
        x == 2,
        x == 2,
        (x < 5) & ((x == 7) | (x == 2)),
        (x == 7) | ((x == 2) | (x == 8)),
        (x < 5) & ((x == 2) | (x == 8) | (x == 1)),
    ]
    for q, m in zip(KAT_CONDITIONS, masks):
        got, _ = oracle.select(db, q)
        np.testing.assert_array_equal(got["col0"], x[m], err_msg=q)


# -------------------------------------------------------------- brute force
@pytest.mark.parametrize("k", range(11))
def test_brute_force(table, k):
    db, c = table
    sql, mask, names = brute_force(c)[k]
    got, ill = oracle.select(db, sql)
    assert ill == 0
    want = {n: c[n][mask] for n in names}
    assert_same(got, want, sql)


@pytest.mark.parametrize("t", [0.1, 10.0, 50.0, 90.0])
def test_config2_brute_force(table, t):
    """The math projection: float32 multiply then add, int32 subtract,
    float32 add then divide, each rounded on its own (no FMA)."""
    db, c = table
    m = c["uniformf"] < f32(t)
    uf, f5, f20 = c["uniformf"][m], c["normalf5"][m], c["normalf20"][m]
    want = {"a": (uf * f32(2.0)) + f5,
            "b": c["uniformi"][m] - c["normali20"][m],
            "c": uf / (f20 + f32(1000.0))}
    got, ill = oracle.select(db, config2_sql(t))
    assert ill == 0
    assert_same(got, want, config2_sql(t))


def test_empty_result(table):
    db, c = table
    got, ill = oracle.select(db, "select id from test where uniformi > 1000")
    assert ill == 0 and len(got["id"]) == 0


# --------------------------------------------------------------- invariants
@pytest.mark.parametrize("sql", [CONFIG1_SQL, QUERY_SUITE[8], config2_sql(50.0)])
def test_threads_ordered_equal_single(table, sql):
    db, _ = table
    one, _ = oracle.select(db, sql)
    many, _ = oracle.select(db, sql, threads=4, ordered=True)
    assert_same(many, one, sql)


def test_unordered_is_same_multiset(table):
    """The reference's multi-core VM reserves output per 64-row block under
    a mutex (virginia.c:488-512): same rows, order not fixed."""
    db, _ = table
    one, _ = oracle.select(db, CONFIG1_SQL)
    many, _ = oracle.select(db, CONFIG1_SQL, threads=4, ordered=False)
    order = np.argsort(many["id"], kind="stable")
    assert_same({k: v[order] for k, v in many.items()}, one, "unordered")


@pytest.mark.parametrize("tablet_rows", [0, 1000, 1 << 20])
def test_tablet_size_invariance(tablet_rows):
    """Same rows whether the table lives in reference 8 MB tablets
    (tablet_rows = 0), many small tablets with ragged tails, or one."""
    db = va.Virginian()
    try:
        db.gen_table("test", 70_001, tablet_rows, seed=3)
        c = gen_columns(70_001, 3)
        _, mask, names = brute_force(c)[10]
        got, _ = oracle.select(db, CONFIG1_SQL)
        assert_same(got, {n: c[n][mask] for n in names}, f"tablet_rows={tablet_rows}")
    finally:
        db.close()


def test_illformed_flagged(db):
    """A program that reads a register no op wrote is flagged (the
    reference would read stale block memory)."""
    simpledb(db, 16)
    O = va
    ops = [(O.OP_Table, 0, 0, 0, None), (O.OP_ResultColumn, O.VIRG_INT, 0, 0, "x"),
           (O.OP_Parallel, 0, 0, 7, None),
           (O.OP_Column, 0, 0, 0, None),
           (O.OP_Gt, 0, 1, 5, 0),              # r1 never written
           (O.OP_Result, 0, 1, 0, None), (O.OP_Converge, 0, 0, 0, None),
           (O.OP_Finish, 0, 0, 0, None)]
    _, ill = oracle.select(db, ops)
    assert ill == 1


def test_type_mismatch_is_an_error(db):
    """int vs float operands abort the reference (REGCMP assert,
    virginia.c:52); the oracle reports an error instead."""
    db.gen_table("test", 100, 0, seed=1)
    with pytest.raises(oracle.OracleError):
        oracle.select(db, "select id from test where uniformi < 2.0")


def test_typed_program_brute_force(db):
    """char / double / int64 arithmetic, Cast, Not/And/Or/Neq keep-branches,
    a masked redefinition and integer division (truncating toward zero),
    written directly in numpy."""
    from tests.helpers import typed_program, typed_table
    c, d, l, i, f = typed_table(db, 5000, 3)
    got, ill = oracle.select(db, typed_program())
    assert ill == 0
    m = i <= 7
    with np.errstate(over="ignore"):
        want = {
            "c0": (c[m].astype(np.int32) * 2).astype(np.int8),
            "c1": d[m] * d[m],
            "c2": l[m],
            "c3": np.where(i[m] == 3, -1, 3).astype(np.int32),
            "c4": np.where(i[m] == 3, -3, np.trunc(i[m] / 3)).astype(np.int32),
            "c5": f[m].astype(np.float64),
            "c6": f[m].astype(np.float64) + d[m],
            "c7": (np.arange(5000, dtype=np.int64) * 3)[m],
        }
    assert_same(got, want, "typed program")
