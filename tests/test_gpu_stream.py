"""Out-of-core path: the table stays in host memory and only the
referenced columns stream through HBM (vg_stream_select, and virg_vm_gpu
with use_stream set as in the reference).  Bit-exact against the oracle in
row order, over batch sizes that split the table at tablet boundaries,
with a ragged last tablet, and over calls in a row that share the
instance's streaming buffers."""
from __future__ import annotations

import pytest

import virginian_amd as va
from oracle import oracle
from tests.helpers import CONFIG1_SQL, QUERY_SUITE, assert_same, config2_sql

pytestmark = pytest.mark.gpu

# the out-of-core benchmark query (right-folded: A and (B or C))
CONFIG3_SQL = ("select id, uniformi, normalf20 from test "
               "where uniformf < 9.24 and normalf5 > 2.0 or normali20 > 10")


@pytest.mark.parametrize("batch_rows", [1, 1 << 20, 3 << 20, 1 << 30])
def test_stream_select(gpu_db, batch_rows):
    gpu_db.gen_table("test", (5 << 20) + 12345, 1 << 20, seed=21)
    for sql in (CONFIG3_SQL, config2_sql(50.0)):
        want, ill = oracle.select(gpu_db, sql, threads=8)
        assert ill == 0
        got, st = gpu_db.stream_select(sql, batch_rows=batch_rows)
        assert st.rows_scanned == (5 << 20) + 12345
        assert st.rows_out == len(next(iter(want.values())))
        assert_same(got, want, f"{sql} batch_rows={batch_rows}")


def test_stream_link_bytes(gpu_db):
    """Only referenced columns cross the link: 28 B/row for the
    out-of-core query (id 8 + five 4-byte columns)."""
    rows = (2 << 20) + 7
    gpu_db.gen_table("test", rows, 1 << 20, seed=22)
    got, st = gpu_db.stream_select(CONFIG3_SQL)
    assert st.h2d_bytes == rows * 28
    assert st.d2h_bytes == st.rows_out * 16
    assert 0.04 < st.rows_out / rows < 0.06


def test_use_stream_boundary(gpu_db):
    """virg_vm_gpu with use_stream: results land in the host result
    tablets (possibly several) in row order."""
    gpu_db.gen_table("test", (3 << 20) + 99, 1 << 20, seed=23)
    gpu_db.use_stream = True
    for sql in (CONFIG1_SQL, CONFIG3_SQL):
        got = gpu_db.query(sql)
        want, _ = oracle.select(gpu_db, sql, threads=8)
        assert_same(got, want, sql)
    gpu_db.use_stream = False


def test_use_mmap_boundary(gpu_db):
    """virg_vm_gpu with use_mmap (the reference's mapped mode, exercised by
    its ConditionsGPUMapped test, src/test/sql.cc:316-330): the data stays
    in host memory; every query of the published suite and the benchmark
    query give the oracle's rows, in order, in the host result tablets."""
    gpu_db.gen_table("test", (2 << 20) + 77, 1 << 19, seed=29)
    gpu_db.use_mmap = True
    try:
        for sql in [CONFIG1_SQL, CONFIG3_SQL] + QUERY_SUITE:
            got = gpu_db.query(sql)
            want, _ = oracle.select(gpu_db, sql, threads=8)
            assert_same(got, want, sql)
    finally:
        gpu_db.use_mmap = False


def test_stream_capacity_error(gpu_db):
    gpu_db.gen_table("test", 100_000, 1 << 15, seed=24)
    with pytest.raises(va.VirginianError):
        gpu_db.stream_select("select id from test where uniformi < 50", capacity=10)


def test_stream_pageable_tablets(gpu_db):
    """Tablets beyond the pin limit are pageable and go through host-thread
    staging; batches mixing pinned and pageable tablets give the same
    result."""
    gpu_db.pin_limit(3 * (40 << 20))
    gpu_db.gen_table("test", (6 << 20) + 5, 1 << 20, seed=25)
    pinned, pageable = gpu_db.host_memory()
    assert pinned > 0 and pageable > 0
    want, _ = oracle.select(gpu_db, CONFIG3_SQL, threads=8)
    for batch_rows in (1 << 20, 5 << 20):
        got, st = gpu_db.stream_select(CONFIG3_SQL, batch_rows=batch_rows)
        assert st.h2d_bytes == ((6 << 20) + 5) * 28
        assert_same(got, want, f"pageable batch_rows={batch_rows}")


def test_stream_buffers_kept_between_calls(gpu_db):
    """The streaming buffers outlive a call: calls on one instance whose
    batches grow (1M rows, the whole table, 1M again), under programs of
    different column and result widths, a call that fails (its capacity is
    too small, which frees them) and the calls after it, through both
    vg_stream_select and the drop-in with use_mmap, all give the oracle's
    rows."""
    gpu_db.gen_table("test", (4 << 20) + 321, 1 << 20, seed=26)
    sqls = [CONFIG3_SQL, "select id from test where uniformi < 30", config2_sql(70.0)]
    wants = {sql: oracle.select(gpu_db, sql, threads=8)[0] for sql in sqls}
    for batch_rows in (1 << 20, 1 << 30, 1 << 20):
        for sql in sqls:
            got, _ = gpu_db.stream_select(sql, batch_rows=batch_rows)
            assert_same(got, wants[sql], f"{sql} batch_rows={batch_rows}")
    with pytest.raises(va.VirginianError):
        gpu_db.stream_select(sqls[1], capacity=10)
    got, _ = gpu_db.stream_select(sqls[2], batch_rows=1 << 30)
    assert_same(got, wants[sqls[2]], "after a failed call")
    gpu_db.use_mmap = True
    try:
        for sql in sqls:
            assert_same(gpu_db.query(sql), wants[sql], f"use_mmap {sql}")
    finally:
        gpu_db.use_mmap = False
