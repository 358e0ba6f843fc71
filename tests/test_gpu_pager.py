"""Queries on the GPU over a table several times the slot pager's limit
(11 reference-sized tablets, 3 kept in host memory, the rest in the
database file): through the drop-in boundary (virg_vm_execute ->
virg_vm_gpu, whose walk unlocks each tablet as it moves on, so earlier
tablets are written out while later ones are read back), as a prepared
query, and streamed from host memory (every tablet kept locked while it
streams: a limit below the table's tablets is refused with a message).
Results equal the oracle's over a twin database kept in memory."""
from __future__ import annotations

import pytest

import virginian_amd as va
from oracle import oracle
from tests.helpers import assert_same
from tests.test_dbfile import SQLS
from tests.test_pager import LIMIT, ROWS, SEED

pytestmark = pytest.mark.gpu


@pytest.fixture
def paged(tmp_path, gpu_db):
    gpu_db.create_db(tmp_path / "paged.db")
    gpu_db.tablet_limit(LIMIT)
    gpu_db.gen_table("test", ROWS, 0, seed=SEED)
    twin = va.Virginian()
    twin.gen_table("test", ROWS, 0, seed=SEED)
    want = [oracle.select(twin, q)[0] for q in SQLS]
    twin.close()
    assert gpu_db.pager_stats()["evictions"] >= 8
    yield gpu_db, want


def test_drop_in(paged):
    db, want = paged
    for q, w in zip(SQLS, want):
        loads = db.pager_stats()["file_loads"]
        assert_same(db.query(q), w, q)
        assert db.pager_stats()["file_loads"] > loads, "tablets came back from the file"
        assert db.pager_stats()["resident"] <= LIMIT + 2     # the last data and result tablets


def test_prepared(paged):
    db, want = paged
    for q, w in zip(SQLS, want):
        p = db.prepare(q)
        try:
            p.run()
            assert_same(p.fetch(), w, q)
        finally:
            p.close()


def test_streamed(paged):
    db, want = paged
    with pytest.raises(va.VirginianError):
        db.stream_select(SQLS[0])
    db.tablet_limit(16)
    got, _ = db.stream_select(SQLS[0])
    assert_same(got, want[0], SQLS[0])
