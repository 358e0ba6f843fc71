"""The slot pager (reference: src/db/findslot.c:19-63, src/db/load.c:22-98):
with a database file open and more tablets than the resident limit, an
unlocked tablet is written to the file, chosen round-robin, and its slot
reused; a tablet not in memory is read back when next loaded; locked
tablets stay put, and when every resident tablet is locked a load fails.
Tables many times the limit keep every byte (against a twin database in
memory), survive close and reopen, and the reference's own VM reads the
file the pager wrote and gives the oracle's results."""
from __future__ import annotations

import ctypes as C

import pytest

import virginian_amd as va
from oracle import oracle, ref
from tests.helpers import assert_same
from tests.test_dbfile import SEED, SQLS, _ref_run

ROWS = 10 * 209_696 + 12_345          # 11 reference-sized (8 MB) tablets of 209,696 rows
LIMIT = 3


def _contents(m) -> bytes:
    """A tablet's keys and column values for its rows (what a query reads;
    the rest of a slot buffer is unused capacity)."""
    base = C.addressof(m)
    parts = [C.string_at(base + m.key_block, m.rows * m.key_stride)]
    for c in range(m.fixed_columns):
        parts.append(C.string_at(base + m.fixed_block + m.fixed_offset[c], m.rows * m.fixed_stride[c]))
    return b"".join(parts)


def _walk(db, tid):
    """(id, rows, keys and values) of every tablet, loading each under a
    lock and letting it go before the next (as the reference's loadnext
    walk does)."""
    out = []
    p = C.POINTER(va.TabletMeta)()
    cur = db.state.db.first_tablet[tid]
    while True:
        assert va.lib.virg_db_load(db.ptr, cur, C.byref(p)) == va.VIRG_SUCCESS
        m = p.contents
        out.append((m.id, m.rows, _contents(m)))
        last, nxt = m.last_tablet, m.next
        assert va.lib.virg_tablet_unlock(db.ptr, m.id) == va.VIRG_SUCCESS
        if last:
            return out
        cur = nxt


@pytest.fixture
def paged(tmp_path):
    db = va.Virginian()
    db.create_db(tmp_path / "paged.db")
    assert db.tablet_limit(LIMIT) == va.VIRG_MEM_TABLETS
    tid = db.gen_table("test", ROWS, 0, seed=SEED)
    twin = va.Virginian()
    ttid = twin.gen_table("test", ROWS, 0, seed=SEED)
    yield db, tid, twin, ttid, tmp_path / "paged.db"
    twin.close()
    db.close()


def test_tables_larger_than_the_limit(paged):
    db, tid, twin, ttid, _ = paged
    st = db.pager_stats()
    assert st["resident"] <= LIMIT
    assert st["evictions"] >= 8
    want = _walk(twin, ttid)
    assert len(want) >= 11
    got = _walk(db, tid)
    assert got == want
    st2 = db.pager_stats()
    assert st2["file_loads"] - st["file_loads"] >= 8, "the walk read tablets back from the file"
    assert st2["resident"] <= LIMIT


def test_locked_tablets_stay_and_all_locked_fails(paged):
    db, tid, twin, ttid, _ = paged
    first = db.state.db.first_tablet[tid]
    p = C.POINTER(va.TabletMeta)()
    assert va.lib.virg_db_load(db.ptr, first, C.byref(p)) == va.VIRG_SUCCESS     # held
    _walk(db, tid)                                           # pages every other tablet through
    assert db.pager_stats(first)["tablet_resident"]
    ids = [t[0] for t in _walk(twin, ttid)]
    q = C.POINTER(va.TabletMeta)()
    held = [first]
    for i in ids[1:LIMIT]:
        assert va.lib.virg_db_load(db.ptr, i, C.byref(q)) == va.VIRG_SUCCESS
        held.append(i)
    # every resident tablet is locked: the next load has nowhere to go
    assert va.lib.virg_db_load(db.ptr, ids[-1], C.byref(q)) == va.VIRG_FAIL
    for i in held:
        va.lib.virg_tablet_unlock(db.ptr, i)
    assert va.lib.virg_db_load(db.ptr, ids[-1], C.byref(q)) == va.VIRG_SUCCESS
    va.lib.virg_tablet_unlock(db.ptr, ids[-1])


def test_round_robin_victims(paged):
    """Each load of a tablet not in memory takes the first slot holding an
    unlocked tablet from the round-robin position on, and moves the
    position past it (reference: src/db/findslot.c:36-57)."""
    db, tid, twin, ttid, _ = paged
    ids = [t[0] for t in _walk(twin, ttid)]
    st = db.state
    p = C.POINTER(va.TabletMeta)()
    moves = 0
    for i in ids * 2:
        if db.pager_stats(i)["tablet_resident"]:
            continue
        status = [st.tablet_slot_status[k] for k in range(va.VIRG_MEM_TABLETS)]
        start = st.tablet_slot_counter
        victim = next(k % va.VIRG_MEM_TABLETS for k in range(start, start + va.VIRG_MEM_TABLETS)
                      if status[k % va.VIRG_MEM_TABLETS] == 1)
        gone = st.tablet_slot_ids[victim]
        assert va.lib.virg_db_load(db.ptr, i, C.byref(p)) == va.VIRG_SUCCESS
        assert st.tablet_slot_ids[victim] == i and not db.pager_stats(gone)["tablet_resident"]
        assert st.tablet_slot_counter == (victim + 1) % va.VIRG_MEM_TABLETS
        va.lib.virg_tablet_unlock(db.ptr, i)
        moves += 1
    assert moves >= len(ids)


def test_close_reopen_and_reference_reads(paged):
    db, tid, twin, ttid, path = paged
    want_images = _walk(twin, ttid)
    want = [oracle.select(twin, q)[0] for q in SQLS]
    db.close_db()
    db2 = va.Virginian()
    try:
        db2.open_db(path)
        db2.tablet_limit(LIMIT)
        assert _walk(db2, db2.table_id("test")) == want_images
        assert db2.pager_stats()["resident"] <= LIMIT
    finally:
        db2.close()
    if ref.available():
        # the reference opens the file the pager wrote; its CPU VM gives the oracle's results
        schema = va.Virginian()
        schema.gen_table("test", 10, 0, seed=SEED)          # same schema, for compiling the programs
        r = ref.RefDB(path=path, existing=True)
        try:
            for q, w in zip(SQLS, want):
                assert_same(_ref_run(r, schema, q, cap=ROWS), w, q)
        finally:
            r.close()
            schema.close()


def test_limit_off_without_file():
    """Without a database file nothing is evicted, whatever the limit."""
    db = va.Virginian()
    try:
        db.tablet_limit(2)
        tid = db.gen_table("test", 5 * 209_716, 0, seed=SEED)
        assert db.pager_stats()["evictions"] == 0
        assert db.pager_stats()["resident"] >= 5
        got, _ = oracle.select(db, SQLS[0])
        assert len(got["id"]) > 0
    finally:
        db.close()
