"""Queries on the GPU over databases opened from files in the reference's
format: tablets are read from the file by virg_db_load on first use, then
run through the drop-in boundary (virg_vm_execute -> virg_vm_gpu) and as a
resident query (vg_query_*); results equal the oracle's over the database
before it was written.  A file written by the reference's own code
(tests/golden/ref_db_20000.db.gz, made in the development container by
tests/golden/make_ref_db.py; the reference never runs here) is queried the
same way."""
from __future__ import annotations

import gzip
from pathlib import Path

import pytest

import virginian_amd as va
from oracle import oracle
from tests.helpers import CONFIG1_SQL, assert_same, config2_sql
from tests.test_oracle import gen_columns

REF_DB = Path(__file__).resolve().parent / "golden" / "ref_db_20000.db.gz"

pytestmark = pytest.mark.gpu

ROWS = 250_000
SEED = 7
SQLS = [CONFIG1_SQL, config2_sql(10.0), config2_sql(90.0),
        "select id, normali5 from test where normali20 > 30 or uniformf < 1.0"]
COLS = ("uniformi", "normali5", "normali20", "uniformf", "normalf5", "normalf20")


def _open(path):
    db = va.Virginian()
    if db.device()[1] == 0:
        db.close()
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X box")
    db.open_db(path)
    return db


def _check_gpu(db, want):
    for q, w in zip(SQLS, want):
        print(f"file query: {q}", flush=True)
        assert_same(db.query(q), w, q)
        dq = db.prepare(q)
        try:
            dq.run()
            assert_same(dq.fetch(), w, q + " (resident)")
        finally:
            dq.close()


def test_gpu_queries_over_our_file(tmp_path):
    path = tmp_path / "ours.db"
    src = va.Virginian()
    src.create_db(path)
    src.gen_table("test", ROWS, 0, seed=SEED)
    want = [oracle.select(src, q)[0] for q in SQLS]
    src.close_db()
    src.close()
    db = _open(path)
    try:
        _check_gpu(db, want)
    finally:
        db.close()


def test_gpu_queries_over_reference_file(tmp_path):
    """The committed file the reference wrote: 20,000 rows of the seeded
    generator (seed 7), which the oracle also reads from the file."""
    path = tmp_path / "ref.db"
    path.write_bytes(gzip.decompress(REF_DB.read_bytes()))
    db = _open(path)
    try:
        tabs = db.tablets(db.table_id("test"))
        assert sum(t.contents.rows for t in tabs) == 20_000
        c = gen_columns(20_000, SEED)
        got = db.query("select id, uniformi, normalf20 from test where uniformi < 1000")
        assert (got["id"] == c["id"]).all() and (got["uniformi"] == c["uniformi"]).all()
        assert (got["normalf20"].view("u4") == c["normalf20"].view("u4")).all()
        want = [oracle.select(db, q)[0] for q in SQLS]
        _check_gpu(db, want)
    finally:
        db.close()
