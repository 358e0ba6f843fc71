"""Database files in the reference's on-disk format (virg_db_create, _open,
_close, _write, _clear; reference src/db/*.c).

Files this library writes open in the reference's own code (oracle/_ref,
built from the reference sources) and files the reference writes open
here, with the same query results; our own round trip keeps every tablet
byte; growing the tablet-info table past its first block (which moves the
first tablet behind the last) keeps every tablet readable on both sides;
and the error paths: a tablet larger than a disk slot, a file that exists,
a missing or truncated file, opening after tables are in memory."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import virginian_amd as va
from oracle import oracle, ref
from tests.helpers import CONFIG1_SQL, assert_same, config2_sql
from tests.test_oracle import gen_columns

ROWS = 250_000          # two reference-sized (8 MB) tablets of the 7-column schema
SEED = 5
SQLS = [CONFIG1_SQL, config2_sql(10.0), config2_sql(90.0),
        "select id, normali5 from test where normali20 > 30 or uniformf < 1.0"]
COLS = ("uniformi", "normali5", "normali20", "uniformf", "normalf5", "normalf20")

needs_ref = pytest.mark.skipif(not ref.available(), reason="reference VM not built (oracle/build_ref.sh)")


def _select(db, sql):
    return oracle.select(db, sql)[0]


def _tablet_images(db, tid) -> list[tuple[int, int, bytes]]:
    """(id, rows, bytes after the meta block) of every tablet of a table."""
    meta = C.sizeof(va.TabletMeta)
    out = []
    for t in db.tablets(tid):
        m = t.contents
        out.append((m.id, m.rows, C.string_at(C.addressof(m) + meta, m.size - meta)))
    return out


def _ref_run(r, db, sql, cap=ROWS):
    """The reference's CPU VM over its database, on the op program this
    library compiles for `sql` (against `db`'s schema)."""
    vm = db.compile(sql)
    try:
        names, types = va.result_schema(vm)
        return r.run(va.ops_array(vm), names, types, cap)
    finally:
        db.free_vm(vm)


def _check_reference_reads(path, want):
    """The reference opens the file and its CPU VM gives `want`."""
    schema = va.Virginian()
    schema.gen_table("test", 10, 0, seed=SEED)         # same schema, for compiling the programs
    r = ref.RefDB(path=path, existing=True)
    try:
        for q, w in zip(SQLS, want):
            assert_same(_ref_run(r, schema, q), w, q)
    finally:
        r.close()
        schema.close()


def test_own_round_trip(tmp_path):
    path = tmp_path / "own.db"
    db = va.Virginian()
    db.create_db(path)
    tid = db.gen_table("test", ROWS, 0, seed=SEED)
    sid = db.create_table("small", va.VIRG_INT, [("a", va.VIRG_INT), ("b", va.VIRG_FLOAT)])
    k = np.arange(1000, dtype=np.int32)
    db.insert(sid, k, [k * 3, (k / 7).astype(np.float32)])
    want = [_select(db, q) for q in SQLS] + [_select(db, "select a, b from small where a > 10")]
    images = _tablet_images(db, tid)
    assert len(images) >= 2
    db.close_db()
    db.close()

    db2 = va.Virginian()
    db2.open_db(path)
    got = [_select(db2, q) for q in SQLS] + [_select(db2, "select a, b from small where a > 10")]
    for g, w, q in zip(got, want, SQLS + ["small"]):
        assert_same(g, w, q)
    assert _tablet_images(db2, db2.table_id("test")) == images
    db2.close()


@needs_ref
def test_reference_opens_our_file(tmp_path):
    path = tmp_path / "ours.db"
    db = va.Virginian()
    db.create_db(path)
    db.gen_table("test", ROWS, 0, seed=SEED)
    want = [_select(db, q) for q in SQLS]
    db.close_db()
    db.close()
    _check_reference_reads(path, want)


@needs_ref
def test_we_open_reference_file(tmp_path):
    c = gen_columns(ROWS, SEED)
    path = tmp_path / "ref.db"
    r = ref.RefDB(path=path)
    r.table("test", c["id"], [(n, c[n]) for n in COLS])
    schema = va.Virginian()
    schema.gen_table("test", 10, 0, seed=SEED)         # same schema, for compiling the programs
    want = [_ref_run(r, schema, q) for q in SQLS]
    r.write_close()
    schema.close()

    db = va.Virginian()
    db.open_db(path)
    tid = db.table_id("test")
    assert db.numrows(tid) == ROWS
    for q, w in zip(SQLS, want):
        assert_same(_select(db, q), w, q)
    db.close()


@needs_ref
def test_tablet_info_growth(tmp_path):
    """60 small tablets: the 17th outgrows the first info block (the first
    tablet moves behind the last), the 49th grows it again in place."""
    path = tmp_path / "many.db"
    db = va.Virginian()
    db.create_db(path)
    tid = db.gen_table("test", 60_000, 1000, seed=SEED)
    assert len(db.tablets(tid)) == 60
    want = [_select(db, q) for q in SQLS]
    images = _tablet_images(db, tid)
    db.close_db()

    db2 = va.Virginian()
    db2.open_db(path)
    assert db2.state.db.alloced_tablets == 16 + 2 * 32
    assert db2.state.db.block_size == va.struct_sizes()[11] + 16 * va.struct_sizes()[6] + va.VIRG_TABLET_SIZE
    assert _tablet_images(db2, db2.table_id("test")) == images
    for q, w in zip(SQLS, want):
        assert_same(_select(db2, q), w, q)
    db2.close()
    db.close()
    _check_reference_reads(path, want)


@needs_ref
def test_growth_with_written_tablets_in_memory(tmp_path):
    """Tablets written with virg_db_write stay in memory with their disk
    entries; when the 17th outgrows the first info block, the entries of
    the resident ones move with the table (the first behind the last) and
    later writes still land in each tablet's own disk slot."""
    path = tmp_path / "resident.db"
    db = va.Virginian()
    db.create_db(path)
    tid = db.gen_table("test", 40_000, 1000, seed=SEED)
    st = db.state
    slots = [i for i in range(va.VIRG_MEM_TABLETS) if st.tablet_slot_status[i] != 0]
    assert len(slots) == 40
    for i in slots:
        assert va.lib.virg_db_write(db.ptr, i) == va.VIRG_SUCCESS
    assert st.db.alloced_tablets == 16 + 32
    want = [_select(db, q) for q in SQLS]
    images = _tablet_images(db, tid)
    db.close_db()                    # writes each again, in place
    db.close()

    db2 = va.Virginian()
    db2.open_db(path)
    assert _tablet_images(db2, db2.table_id("test")) == images
    for q, w in zip(SQLS, want):
        assert_same(_select(db2, q), w, q)
    db2.close()
    _check_reference_reads(path, want)


@needs_ref
def test_reference_file_past_first_info_block(tmp_path):
    """The reference writes 18 tablets of 8 MB: its 17th moves the first
    tablet behind the last (src/db/write.c:105-170); every tablet reads
    back here."""
    n = 17 * 209_716 + 1000
    c = gen_columns(n, SEED)
    path = tmp_path / "refmany.db"
    r = ref.RefDB(path=path)
    r.table("test", c["id"], [(k, c[k]) for k in COLS])
    schema = va.Virginian()
    schema.gen_table("test", 10, 0, seed=SEED)
    want = [_ref_run(r, schema, q, cap=n) for q in SQLS]
    r.write_close()
    schema.close()

    db = va.Virginian()
    db.open_db(path)
    assert db.state.db.alloced_tablets == 16 + 32
    tid = db.table_id("test")
    assert db.numrows(tid) == n and len(db.tablets(tid)) == 18
    for q, w in zip(SQLS, want):
        assert_same(_select(db, q), w, q)
    db.close()


def test_errors(tmp_path, capfd):
    path = tmp_path / "e.db"
    db = va.Virginian()
    db.create_db(path)
    db.close_db()
    with pytest.raises(va.VirginianError):
        db.create_db(path)                              # exists (O_EXCL, as the reference)
    with pytest.raises(va.VirginianError):
        db.open_db(tmp_path / "missing.db")
    db.gen_table("test", 100, 0, seed=1)
    with pytest.raises(va.VirginianError):
        db.open_db(path)                                # tables are already in memory
    db.close()

    big = va.Virginian()
    big.create_db(tmp_path / "big.db")
    big.gen_table("test", 300_000, 300_000, seed=1)     # one 12 MB tablet: no disk slot holds it
    with pytest.raises(va.VirginianError):
        big.close_db()
    assert "larger than a database file's disk slot" in capfd.readouterr().err
    big.close()

    trunc = tmp_path / "trunc.db"
    trunc.write_bytes(b"\0" * 100)
    t = va.Virginian()
    with pytest.raises(va.VirginianError):
        t.open_db(trunc)
    t.close()
