"""GPU parity of the specialised single-pass kernels (one compiled per
program structure at prepare time): against the CPU oracle, bit for bit
and in row order, over the benchmark queries, ragged layouts, typed
programs, random programs, relaunches and graph replays."""
from __future__ import annotations

import numpy as np
import pytest

import virginian_amd as va
from oracle import oracle
from tests.helpers import (CONFIG1_SQL, CONFIG2_THRESHOLDS, KAT_CONDITIONS, QUERY_SUITE, assert_same,
                           cast_programs, config2_sql, simpledb, typed_program, typed_table)
from virginian_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture
def jit():
    # every query, the small tables too (the default leaves tables of fewer
    # than 8192 rows per CU on the interpreter)
    old = va.jit_policy(2)
    yield
    va.jit_policy(old)


def _run(db, prog, capacity=0):
    q = db.prepare(prog, capacity=capacity)
    try:
        k = q.kernel()
        assert k["fused"], "the program has no specialised kernel"
        n = q.run()
        out = q.fetch()
        assert all(len(v) == n for v in out.values())
        return out, k
    finally:
        q.close()


@pytest.mark.parametrize("threshold", CONFIG2_THRESHOLDS)
def test_config2(gpu_db, jit, threshold):
    """Config 2 (math projection, selectivity sweep) over 4M rows in
    1M-row tablets: every chunk of the look-back chain in use."""
    gpu_db.gen_table("test", 4 << 20, 1 << 20, seed=2)
    sql = config2_sql(threshold)
    got, k = _run(gpu_db, sql)
    print(f"T={threshold}: {len(got['a'])} rows, kernel {k}", flush=True)
    want, ill = oracle.select(gpu_db, sql, threads=8)
    assert ill == 0
    assert_same(got, want, sql)


def test_held_launch_bound(gpu_db, jit, monkeypatch):
    """A one-column filter is held in registers, by default in 16K-row
    chunks at the compiler's own occupancy.  With 8K-row chunks and the
    three-wave launch bound (VG_JIT_S=2, VG_JIT_WAVES=3) the benchmark
    query's kernel keeps the bound (a few dwords spilled), while selecting
    all six columns spills past the limit and is rebuilt without it.  All
    are bit exact against the oracle."""
    gpu_db.gen_table("test", 4 << 20, 1 << 20, seed=4)
    six = "select id, uniformf, normalf5, uniformi, normali20, normalf20 from test where uniformf < 50.0"
    cases = [({}, config2_sql(10.0), 0, 4),
             ({"VG_JIT_S": "2", "VG_JIT_WAVES": "3"}, config2_sql(10.0), 3, 2),
             ({"VG_JIT_S": "2", "VG_JIT_WAVES": "3"}, six, 0, 2)]
    for env, sql, waves, tiles in cases:
        for k in ("VG_JIT_S", "VG_JIT_WAVES"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        got, k = _run(gpu_db, sql)
        print(f"{env} {sql}: {len(next(iter(got.values())))} rows, kernel {k}", flush=True)
        assert k["held"] and k["pipelined"] and k["waves"] == waves and k["tiles_per_chunk"] == tiles, k
        want, ill = oracle.select(gpu_db, sql, threads=8)
        assert ill == 0
        assert_same(got, want, sql)


@pytest.mark.parametrize("tablet_rows", [0, 1 << 20])
def test_config1(gpu_db, jit, tablet_rows):
    gpu_db.gen_table("test", 1 << 20, tablet_rows, seed=1)
    got, _ = _run(gpu_db, CONFIG1_SQL)
    want, ill = oracle.select(gpu_db, CONFIG1_SQL)
    assert ill == 0
    assert_same(got, want, CONFIG1_SQL)


def test_kats_and_suite(gpu_db, jit):
    """The reference's condition KATs (src/test/sql.cc:15-37) and its
    10-query suite (src/test/queries.h)."""
    simpledb(gpu_db, 10)
    for q in KAT_CONDITIONS:
        got, _ = _run(gpu_db, q)
        want, ill = oracle.select(gpu_db, q)
        assert ill == 0
        assert_same(got, want, q)


def test_reference_suite_table(gpu_db, jit):
    """The reference's published suite over a table of its own generator's
    distributions (int key, truncated ints, src/test/queries.h over
    db/generate.c's table) in 300K-row tablets (about the reference's 8 MB
    tablets, the last one ragged)."""
    from virginian_amd.workload import suite_table
    suite_table(gpu_db, 1 << 20, 300_000, seed=5)
    for q in QUERY_SUITE:
        got, _ = _run(gpu_db, q)
        want, ill = oracle.select(gpu_db, q, threads=8)
        assert ill == 0, q
        assert_same(got, want, q)


def test_query_suite(gpu_db, jit):
    gpu_db.gen_table("test", 1 << 20, 1 << 18, seed=3)
    for q in QUERY_SUITE:
        got, _ = _run(gpu_db, q)
        want, ill = oracle.select(gpu_db, q)
        assert ill == 0, q
        assert_same(got, want, q)


@pytest.mark.parametrize("rows,tablet_rows", [(1, 0), (63, 64), (1000, 64), (4095, 4000), (4097, 4096),
                                              (16385, 5000), (100003, 1 << 15), (300001, 0)])
def test_layouts(gpu_db, jit, rows, tablet_rows):
    """Partial tiles, chunks that span tablets, tablets smaller than a
    tile, a ragged last tablet."""
    gpu_db.gen_table("test", rows, tablet_rows, seed=rows)
    for sql in (CONFIG1_SQL, config2_sql(50.0), "select id from test where uniformi < 101"):
        got, _ = _run(gpu_db, sql)
        want, ill = oracle.select(gpu_db, sql)
        assert ill == 0
        assert_same(got, want, f"{sql} rows={rows} tablet_rows={tablet_rows}")


def test_empty_and_full(gpu_db, jit):
    gpu_db.gen_table("test", 70000, 1 << 14, seed=5)
    got, _ = _run(gpu_db, "select id, uniformf from test where uniformf < 0.0")
    assert len(got["id"]) == 0
    got, _ = _run(gpu_db, "select id, normali5 from test where uniformf < 1000.0")
    assert len(got["id"]) == 70000
    assert np.array_equal(got["id"], np.arange(70000))


def test_typed_and_unaligned(gpu_db, jit):
    """char / double / int64 columns (a char column first, so the other
    column bases are not 16-byte aligned: the per-row load variant)."""
    typed_table(gpu_db, 5000, 3)
    for ops in [typed_program()] + cast_programs():
        want, ill = oracle.select(gpu_db, ops)
        assert ill == 0
        got, _ = _run(gpu_db, ops)
        assert_same(got, want, "typed program")


def test_relaunch_and_alternate(gpu_db, jit):
    """The chunk counter and epoch carry over between launches; queries
    sharing a kernel keep their own state."""
    gpu_db.gen_table("test", 1 << 20, 1 << 18, seed=9)
    qs = [gpu_db.prepare(config2_sql(t)) for t in (10.0, 60.0)]
    wants = [oracle.select(gpu_db, config2_sql(t))[0] for t in (10.0, 60.0)]
    for it in range(6):
        q, want = qs[it % 2], wants[it % 2]
        q.run()
        assert_same(q.fetch(), want, f"launch {it}")
    for q in qs:
        for _ in range(3):
            va.lib.vg_query_launch(q.h)
    for q, want in zip(qs, wants):
        q.sync()
        assert_same(q.fetch(), want, "after back-to-back launches")
        q.close()


def test_graph_replays(gpu_db, jit):
    """Specialised launches captured in a HIP graph replay with the right
    result every time (nothing on the host moves between replays)."""
    gpu_db.gen_table("test", 600_000, 1 << 17, seed=12)
    sqls = [config2_sql(25.0), CONFIG1_SQL]
    qs = [gpu_db.prepare(s) for s in sqls]
    wants = [oracle.select(gpu_db, s)[0] for s in sqls]
    b = va.GraphBatch(qs)
    for rep in range(4):
        rows = b.run()
        for q, want, n in zip(qs, wants, rows):
            q.last_rows = n
            assert_same(q.fetch(), want, f"replay {rep}")
        qs[0].run()          # eager launch between replays
    b.close()
    for q in qs:
        q.close()


def test_capacity_overflow(gpu_db, jit):
    """Rows past the capacity are dropped and the query reports it."""
    gpu_db.gen_table("test", 200_000, 1 << 16, seed=4)
    q = gpu_db.prepare("select id from test where uniformi < 50", capacity=1000)
    assert q.kernel()["fused"]
    with pytest.raises(va.VirginianError):
        q.run()
    q.close()
    q = gpu_db.prepare("select id from test where uniformi < 50")
    n = q.run()
    assert n == len(oracle.select(gpu_db, "select id from test where uniformi < 50")[0]["id"])
    q.close()


@pytest.mark.parametrize("tablet_rows", [0, 1 << 16])
def test_random_programs(gpu_db, jit, tablet_rows):
    """Seeded random programs of up to 40 ops (OR groups with waiting
    rows, masked redefinitions, Invalid, Rowid, float division)."""
    gpu_db.gen_table("test", 300_001, tablet_rows, seed=11)
    rng = np.random.default_rng(177 + tablet_rows)
    for i in range(20):
        sel = float(rng.uniform(0.0001, 0.8)) if rng.random() < 0.7 else None
        prog = W.random_program(rng, max_ops=40, selectivity=sel)
        print(f"program {i}: {len(prog)} ops", flush=True)
        want, ill = oracle.select(gpu_db, prog)
        assert ill == 0
        got, _ = _run(gpu_db, prog)
        assert_same(got, want, f"program {i}: {prog}")


@pytest.mark.parametrize("stage", ["1", "0"], ids=["staged", "registers"])
def test_mask_logic_regression(gpu_db, jit, monkeypatch, stage):
    """Filters that drop the rows failing two OR-ed compares, with the
    filter columns staged in LDS or loaded into registers: compiled with
    v_bitop3 the mask logic once got a wrong truth table (0 rows instead
    of 509 for P187; 2210 instead of 3244 for the seeded program)."""
    from tests.test_kernel_cpu import P187
    monkeypatch.setenv("VG_JIT_STAGE", stage)
    gpu_db.gen_table("test", 4096, 0, seed=11)
    rng = np.random.default_rng(177)
    for i in range(18):
        sel = float(rng.uniform(0.0001, 0.8)) if rng.random() < 0.7 else None
        prog = W.random_program(rng, max_ops=40, selectivity=sel)
    for prog in (P187, prog):
        want, ill = oracle.select(gpu_db, prog)
        assert ill == 0
        got, k = _run(gpu_db, prog)
        assert k["staged"] == (stage == "1")
        assert_same(got, want, f"{prog}")


def test_small_tables_stay_on_the_interpreter(gpu_db):
    """Under the default policy a table of fewer than 8192 rows per CU runs
    on the interpreter (its chunks could not fill the device); policy 2
    takes the specialised kernel for it too; a larger table gets it by
    default.  The results are the oracle's either way."""
    name, cus = gpu_db.device()
    gpu_db.gen_table("small", 1 << 20, 1 << 20, seed=7)
    gpu_db.gen_table("test", 4 << 20, 1 << 20, seed=7)
    assert (1 << 20) < cus * 8192 <= (4 << 20), cus
    small_sql = CONFIG1_SQL.replace("from test", "from small")
    old = va.jit_policy(1)
    try:
        for policy, sql, fused in ((1, small_sql, False), (2, small_sql, True), (1, CONFIG1_SQL, True)):
            va.jit_policy(policy)
            q = gpu_db.prepare(sql)
            try:
                assert q.kernel()["fused"] == fused, (policy, sql)
                q.run()
                want, ill = oracle.select(gpu_db, sql)
                assert ill == 0
                assert_same(q.fetch(), want, f"policy {policy}: {sql}")
            finally:
                q.close()
    finally:
        va.jit_policy(old)
