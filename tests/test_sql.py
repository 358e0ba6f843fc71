"""SQL front-end: op-program shape against the reference's own tests.

The parser and code generator are this library's (the reference's grammar
needs bison/flex); these checks pin them to the programs the reference
emits: src/test/sql.cc and the hand trace of generate.c for the config-1
query.  No GPU is needed to compile a query.
"""
from __future__ import annotations

import pytest

import virginian_amd as va
from tests.helpers import CONFIG1_SQL, KAT_CONDITIONS, QUERY_SUITE, config2_sql, simpledb


def _bench_table(db, rows=64):
    db.gen_table("test", rows, 0, seed=1)


def test_basic_shape(db):
    """reference: src/test/sql.cc:40-59 (Basic)."""
    simpledb(db, 1)
    vm = db.compile("select col0 from test")
    try:
        p = va.program(vm)
        assert p[0][0] == va.OP_Table
        assert p[1][0] == va.OP_ResultColumn and p[1][4] == "col0"
        assert p[2][0] == va.OP_Parallel
    finally:
        db.free_vm(vm)


def test_exprops_shape(db):
    """reference: src/test/sql.cc:129-155 (ExprOps): the result column is
    named by the parenthesised expression and 5*20 is folded to 100."""
    simpledb(db, 1)
    vm = db.compile("select col0 + 5 * 20 from test")
    try:
        p = va.program(vm)
        assert p[0][0] == va.OP_Table
        assert p[1][0] == va.OP_ResultColumn and p[1][4] == "(col0+(5*20))"
        assert p[2][0] == va.OP_Parallel
        assert p[4][0] == va.OP_Integer and p[4][2] == 100
    finally:
        db.free_vm(vm)


def test_columns_shape(db):
    """reference: src/test/sql.cc:355-388 (Columns): Result starts at the
    col0 register, covers 2 registers, col0's register below col1's."""
    simpledb(db, 1)
    vm = db.compile("select col0, col1 from test")
    try:
        p = va.program(vm)
        reg = {}
        for op, p1, p2, p3, p4 in p:
            if op == va.OP_Column:
                reg[p2] = p1
        results = [o for o in p if o[0] == va.OP_Result]
        assert len(results) == 1
        assert results[0][1] == reg[0] and results[0][2] == 2
        assert reg[0] < reg[1]
    finally:
        db.free_vm(vm)


def test_config1_program(db):
    """The config-1 program, op for op, as generate.c:591-776 emits it
    (register numbers after the contiguity renumbering, :650-674)."""
    _bench_table(db)
    vm = db.compile(CONFIG1_SQL)
    try:
        p = va.program(vm)
    finally:
        db.free_vm(vm)
    O = va
    want = [
        (O.OP_Table, 0, None, None, None),
        (O.OP_ResultColumn, O.VIRG_INT64, None, None, "id"),
        (O.OP_ResultColumn, O.VIRG_INT, None, None, "uniformi"),
        (O.OP_ResultColumn, O.VIRG_FLOAT, None, None, "normalf5"),
        (O.OP_Parallel, None, None, 16, None),
        (O.OP_Column, 4, 0, None, None),
        (O.OP_Integer, 0, 50, None, None),
        (O.OP_Ge, 4, 0, 14, 0),
        (O.OP_Column, 1, 1, None, None),
        (O.OP_Integer, 2, 0, None, None),
        (O.OP_Gt, 1, 2, 12, 1),
        (O.OP_Invalid, None, None, None, None),
        (O.OP_Rowid, 3, None, None, None),
        (O.OP_Column, 5, 4, None, None),
        (O.OP_Result, 3, 3, None, None),
        (O.OP_Converge, None, None, None, None),
        (O.OP_Finish, None, None, None, None),
    ]
    assert len(p) == len(want)
    for k, (got, exp) in enumerate(zip(p, want)):
        for field, (g, e) in enumerate(zip(got, exp)):
            if e is not None:
                assert g == e, f"op {k} field {field}: {got} vs {exp}"


def test_conditions_right_fold(db):
    """AND/OR chains fold to the right (sql.y:126-149, generate.c:477-578):
    `A or B and C` is `A or (B and C)`; that is what makes the reference's
    second condition select {8, 9} (sql.cc:17, :28)."""
    simpledb(db, 10)
    from oracle import oracle
    got, _ = oracle.select(db, KAT_CONDITIONS[1])
    assert sorted(got["col0"].tolist()) == [8, 9]


@pytest.mark.parametrize("sql", QUERY_SUITE + [CONFIG1_SQL, config2_sql(50.0)])
def test_suite_compiles(db, sql):
    """Every published query compiles within the reference limits: one
    Table, Parallel before the section, one Result right before Converge,
    every jump forward and no further than the Result (Q15 of the
    reference's GPU contract), at most 16 registers."""
    _bench_table(db)
    vm = db.compile(sql)
    try:
        p = va.program(vm)
    finally:
        db.free_vm(vm)
    assert len(p) <= va.VIRG_OPS
    ops = [o[0] for o in p]
    assert ops.count(va.OP_Table) == 1 and ops.count(va.OP_Result) == 1
    par, res = ops.index(va.OP_Parallel), ops.index(va.OP_Result)
    assert ops[res + 1] == va.OP_Converge and ops[-1] == va.OP_Finish
    assert p[par][3] == res + 2
    for k in range(par + 1, res):
        op, p1, p2, p3, p4 = p[k]
        if op in (va.OP_Eq, va.OP_Neq, va.OP_Lt, va.OP_Le, va.OP_Gt, va.OP_Ge):
            assert k < p3 <= res, (k, p[k])
        if op in (va.OP_Column, va.OP_Rowid, va.OP_Integer, va.OP_Float):
            assert 0 <= p1 < va.VIRG_REGS


def test_reject_bad_sql(db):
    """Parse errors and unknown names are reported, not fatal (the
    reference exits the process, sql.y:206-209)."""
    simpledb(db, 1)
    for bad in ["select from test", "select col9 from test", "select col0 from nosuch",
                "select col0 from test where (col0 = 1 or col0 = 2) and col0 = 3",
                "select col0 from test where col0 -5"]:
        with pytest.raises(va.VirginianError):
            vm = db.compile(bad)
            db.free_vm(vm)


def test_constant_fold_quirk(db):
    """Mixed int/float constant folding goes through the int value
    (generate.c:336-337): 1.5 + 1.5 folds to 2.0."""
    _bench_table(db)
    vm = db.compile("select id from test where uniformf < 1.5 + 1.5")
    try:
        p = va.program(vm)
    finally:
        db.free_vm(vm)
    floats = [o[4] for o in p if o[0] == va.OP_Float]
    assert floats == [2.0]
