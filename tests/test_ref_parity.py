"""The oracle against the reference's own CPU VM (oracle/_ref, compiled
from the reference sources by oracle/build_ref.sh), on the same table and
the same programs: the published queries, the reference's condition KATs
and seeded random programs with forward jumps, OR groups, masked
redefinitions and Invalid.  Results must agree bit for bit in the
reference's single-core row order; the reference's multi-core VM must give
the same multiset.  Skipped where the reference sources were not built."""
from __future__ import annotations

import numpy as np
import pytest

import virginian_amd as va
from oracle import oracle, ref
from tests.helpers import (CONFIG1_SQL, KAT_CONDITION_ROWS, KAT_CONDITIONS, QUERY_SUITE,
                           assert_same, config2_sql, simpledb)
from tests.test_oracle import gen_columns
from virginian_amd import workload as W

pytestmark = pytest.mark.skipif(not ref.available(), reason="reference VM not built (oracle/build_ref.sh)")

ROWS = 250_000          # more than one 8 MB reference tablet (about 209k rows of this schema)
SEED = 5


def test_kat_conditions_reference(db):
    """The reference VM run through the harness reproduces src/test/sql.cc's
    condition counts (:15-37, :241-274) and agrees with the oracle.  Runs
    before the module's shared reference database exists (one per process)."""
    r = ref.RefDB()
    try:
        i = np.arange(10, dtype=np.int32)
        r.table("test", i, [("col0", i), ("col1", i + 1), ("col2", i + 2)])
        simpledb(db, 10)
        for q, rows in zip(KAT_CONDITIONS, KAT_CONDITION_ROWS):
            vm = db.compile(q)
            try:
                names, types = va.result_schema(vm)
                got = r.run(va.ops_array(vm), names, types, 10)
            finally:
                db.free_vm(vm)
            assert len(got["col0"]) == rows, q
            want, _ = oracle.select(db, q)
            assert_same(want, got, q)
    finally:
        r.close()


@pytest.fixture(scope="module")
def tables():
    c = gen_columns(ROWS, SEED)
    r = ref.RefDB()
    r.table("test", c["id"], [(n, c[n]) for n in ("uniformi", "normali5", "normali20",
                                                    "uniformf", "normalf5", "normalf20")])
    db = va.Virginian()
    db.gen_table("test", ROWS, 0, seed=SEED)
    yield r, db, c
    db.close()
    r.close()


def _both(r, db, prog):
    if isinstance(prog, str):
        vm = db.compile(prog)
        try:
            ops = va.ops_array(vm)
            names, types = va.result_schema(vm)
            got = r.run(ops, names, types, ROWS)
        finally:
            db.free_vm(vm)
    else:
        arr, n, keep = va.make_ops(prog)
        names = [o[4] for o in prog if o[0] == va.OP_ResultColumn]
        types = [o[1] for o in prog if o[0] == va.OP_ResultColumn]
        got = r.run((arr, n), names, types, ROWS)
    want, ill = oracle.select(db, prog)
    return got, want, ill


@pytest.mark.parametrize("sql", QUERY_SUITE + [CONFIG1_SQL, config2_sql(10.0), config2_sql(90.0)])
def test_published_queries(tables, sql):
    r, db, _ = tables
    got, want, ill = _both(r, db, sql)
    assert ill == 0
    assert len(want[next(iter(want))]) > 0
    assert_same(want, got, sql)


def test_random_programs(tables):
    r, db, _ = tables
    rng = np.random.default_rng(2024)
    nonempty = 0
    for i in range(150):
        sel = float(rng.uniform(0.0001, 0.8)) if rng.random() < 0.7 else None
        prog = W.random_program(rng, max_ops=va.VIRG_OPS, selectivity=sel)
        got, want, ill = _both(r, db, prog)
        assert ill == 0, (i, prog)
        assert_same(want, got, f"program {i}: {prog}")
        nonempty += len(next(iter(want.values()))) > 0
    assert nonempty > 100


def test_multi_core_reference_same_multiset(tables):
    """The reference's multi-core VM (threads take whole tablets, output
    reserved per 64-row block under a mutex) returns the same rows."""
    r, db, c = tables
    r.close()
    rm = ref.RefDB(multi=True, threads=4)
    try:
        rm.table("test", c["id"], [(n, c[n]) for n in ("uniformi", "normali5", "normali20",
                                                        "uniformf", "normalf5", "normalf20")])
        for sql in (CONFIG1_SQL, QUERY_SUITE[8], config2_sql(50.0)):
            got, want, _ = _both(rm, db, sql)
            first = next(iter(got))
            if first == "id":
                order = np.argsort(got["id"], kind="stable")
                got = {k: v[order] for k, v in got.items()}
                assert_same(want, got, sql)
            else:
                for k in want:
                    assert np.array_equal(np.sort(want[k].view(f"u{want[k].dtype.itemsize}")),
                                          np.sort(got[k].view(f"u{got[k].dtype.itemsize}"))), (sql, k)
    finally:
        rm.close()
        # leave a single-core database for any later test of the module
        r.__init__()
        r.table("test", c["id"], [(n, c[n]) for n in ("uniformi", "normali5", "normali20",
                                                       "uniformf", "normalf5", "normalf20")])


def test_cast_reference():
    """Cast, which no query generates, against the reference VM itself on
    the programs where the reference's Cast is well defined
    (tests/helpers.py cast_programs: the op before each Cast carries the
    same type and register, and no value widens or keeps its type)."""
    from tests.helpers import cast_programs, typed_table
    rows = 40_000
    db = va.Virginian()
    r = ref.RefDB()
    try:
        c, d, l, i, f = typed_table(db, rows, 21)
        r.table("test", np.arange(rows, dtype=np.int64) * 3,
                [("c", c), ("d", d), ("l", l), ("i", i), ("f", f)])
        for k, prog in enumerate(cast_programs()):
            arr, n, _ = va.make_ops(prog)
            names = [o[4] for o in prog if o[0] == va.OP_ResultColumn]
            types = [o[1] for o in prog if o[0] == va.OP_ResultColumn]
            got = r.run((arr, n), names, types, rows)
            want, ill = oracle.select(db, prog)
            assert ill == 0
            assert len(want["c0"]) > (rows // 4 if k else rows - 1)
            assert_same(want, got, f"cast program {k}")
    finally:
        r.close()
        db.close()
