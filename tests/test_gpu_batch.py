"""Batched execution: many prepared queries of different programs over
tables of different sizes, run together, give each query exactly the
result it gives alone (and the oracle's, bit for bit and in row order),
on every replay, with eager launches of the same queries in between."""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import pytest

import virginian_amd as va
from oracle import oracle
from tests.helpers import assert_same
from virginian_amd import workload as W

pytestmark = pytest.mark.gpu


def _mix(db, nq, seed, sizes):
    rng = np.random.default_rng(seed)
    tids = []
    for k, rows in enumerate(sizes):
        tablet_rows = int(rng.choice([0, 1 << 14, 1 << 17]))
        db.gen_table(f"t{k}", rows, tablet_rows, seed=seed + k)
        tids.append(db.table_id(f"t{k}"))
    progs = []
    for _ in range(nq):
        tid = int(rng.choice(tids))
        sel = float(rng.uniform(0.0001, 0.8))
        progs.append(W.random_program(rng, max_ops=40, selectivity=sel, table=tid))
    return progs


def test_graph_batch(gpu_db):
    progs = _mix(gpu_db, 48, 5, [1000, 4097, 65_536, 300_001, 1_000_000, 17])
    want = []
    for p in progs:
        w, ill = oracle.select(gpu_db, p)
        assert ill == 0
        want.append(w)
    qs = [gpu_db.prepare(p) for p in progs]
    b = va.GraphBatch(qs)
    try:
        for replay in range(3):
            print(f"replay {replay}", flush=True)
            rows = b.run()
            for i, (q, w) in enumerate(zip(qs, want)):
                assert rows[i] == len(next(iter(w.values()))), f"query {i} replay {replay}"
                assert_same(q.fetch(), w, f"query {i} replay {replay}")
            # eager launches of some queries between replays
            for i in range(0, len(qs), 7):
                qs[i].run()
                assert_same(qs[i].fetch(), want[i], f"query {i} eager")
    finally:
        b.close()
        for q in qs:
            q.close()


def test_graph_batch_rejects_duplicates(gpu_db):
    gpu_db.gen_table("test", 5000, 1 << 12, seed=1)
    q = gpu_db.prepare("select id from test where uniformi < 50")
    try:
        with pytest.raises(va.VirginianError):
            va.GraphBatch([q, q])
    finally:
        q.close()


@pytest.fixture
def mix(gpu_db):
    progs = _mix(gpu_db, 64, 9, [1000, 4097, 65_536, 300_001, 1_000_000, 2_500_000, 17, 200_000])
    want = []
    for p in progs:
        w, ill = oracle.select(gpu_db, p)
        assert ill == 0
        want.append(w)
    qs = [gpu_db.prepare(p) for p in progs]
    yield qs, want
    for q in qs:
        q.close()


def _check(qs, want, rows, what):
    for i, (q, w) in enumerate(zip(qs, want)):
        assert rows[i] == len(next(iter(w.values()))), f"query {i} {what}"
        assert_same(q.fetch(), w, f"query {i} {what}")


@pytest.mark.parametrize("tpi,order,div,split", [(1, 0, 1, 0), (8, 1, 1, 0), (64, 2, 3, 0), (3, 1, 7, 1),
                                                  (16, 0, 2, 1)])
def test_persistent_batch(mix, tpi, order, div, split):
    """Persistent launches over item lists give every query its own
    result, for item sizes, list orders, grids and query groupings that
    split the work differently, on every replay, with eager launches in
    between."""
    qs, want = mix
    b = va.PersistentBatch(qs, tiles_per_item=tpi, order=order)
    try:
        b.configure(tpi, order, div, split)
        cfg = b.config()
        assert cfg["groups"] >= 1 and (split == 0 or cfg["groups"] == 1)
        for replay in range(2):
            print(f"tpi {tpi} order {order} div {div} split {split} groups {cfg['groups']} replay {replay}", flush=True)
            t0 = time.perf_counter()
            rows, lat = b.run()
            wall_ms = (time.perf_counter() - t0) * 1e3
            _check(qs, want, rows, f"tpi {tpi} replay {replay}")
            # every query with tiles has a finish time inside the run
            assert all(0 < x <= wall_ms for q, x in zip(qs, lat) if q.tiles), (wall_ms, lat)
            for i in range(0, len(qs), 9):
                qs[i].run()
                assert_same(qs[i].fetch(), want[i], f"query {i} eager")
    finally:
        b.close()


def test_tuned_batch_and_plain(mix):
    """The autotuner's pick gives the same results, and so do the plain
    per-query launches it is compared with."""
    qs, want = mix
    b = va.PersistentBatch(qs)
    try:
        cfg, ms = b.tune(iters=1)
        print("tuned", cfg, ms, flush=True)
        assert ms > 0
        rows, lat = b.run()
        _check(qs, want, rows, "tuned")
    finally:
        b.close()
    ms, lat = va.plain_time(qs, 1, 2)
    assert ms > 0 and len(lat) == len(qs) and lat[-1] > 0
    rows = [q.sync() for q in qs]
    _check(qs, want, rows, "plain")


def test_persistent_batch_overflow(gpu_db):
    gpu_db.gen_table("test", 100_000, 1 << 15, seed=4)
    q1 = gpu_db.prepare("select id from test where uniformi < 50", capacity=1000)
    q2 = gpu_db.prepare("select id from test where uniformi < 5")
    b = va.PersistentBatch([q1, q2])
    try:
        with pytest.raises(va.VirginianError):
            b.run()
        # the next run of a batch without the overflowing query is fine
    finally:
        b.close()
    b = va.PersistentBatch([q2])
    try:
        rows, _ = b.run()
        want, _ = oracle.select(gpu_db, "select id from test where uniformi < 5")
        assert rows[0] == len(want["id"])
        assert_same(q2.fetch(), want, "after overflow")
    finally:
        b.close()
        q1.close()
        q2.close()


@pytest.mark.parametrize("tpi,split", [(1, 0), (4, 1), (16, 0)])
def test_single_pass(mix, tpi, split):
    """One pass with look-back offsets: every query's result in row order,
    on every replay, when the capacity is large enough."""
    qs, want = mix
    b = va.PersistentBatch(qs, tiles_per_item=tpi, split=split, single=1)
    try:
        for replay in range(2):
            print(f"single tpi {tpi} split {split} replay {replay}", flush=True)
            rows, lat, flags = b.results()
            assert not any(flags)
            _check(qs, want, rows, f"single tpi {tpi} replay {replay}")
            assert max(lat) > 0
    finally:
        b.close()


def test_single_pass_capacity(gpu_db):
    """A result that does not fit is reported, with its exact row count,
    and never written past the capacity; one that fits exactly is whole."""
    gpu_db.gen_table("test", 300_001, 1 << 16, seed=7)
    sqls = ["select id, uniformf from test where uniformi < 30",
            "select normalf5 from test where uniformf < 0.5",
            "select id from test where normali20 > 100"]
    exact = []
    for sql in sqls:
        w, _ = oracle.select(gpu_db, sql)
        exact.append((w, len(next(iter(w.values())))))
    assert exact[2][1] == 0
    qs = [gpu_db.prepare(sqls[0], capacity=exact[0][1]),          # fits exactly
          gpu_db.prepare(sqls[1], capacity=max(exact[1][1] - 1, 1)),  # one short
          gpu_db.prepare(sqls[0], capacity=1000),                  # far short
          gpu_db.prepare(sqls[2], capacity=1)]                     # empty result
    b = va.PersistentBatch(qs, tiles_per_item=2, single=1)
    try:
        rows, _, flags = b.results()
        assert flags == [0, 1, 1, 0]
        assert rows[:3] == [exact[0][1], exact[1][1], exact[0][1]]
        assert rows[3] == 0
        assert_same(qs[0].fetch(), exact[0][0], "exact fit")
    finally:
        b.close()
        for q in qs:
            q.close()


def test_batch_holds_its_queries(gpu_db):
    """While a batch holds a query's launch arguments, the query's outputs
    cannot change and the query cannot be freed (nothing is launched until
    the batch has let go)."""
    gpu_db.gen_table("test", 5000, 1024, seed=1)
    q = gpu_db.prepare("select id from test where uniformi < 50")
    q2 = gpu_db.prepare("select id from test where uniformi < 20")
    buf = gpu_db.device_alloc(5000 * 8)
    ptrs = (C.c_void_p * 1)(buf)
    for make in (lambda: va.GraphBatch([q, q2]), lambda: va.PersistentBatch([q, q2])):
        b = make()
        assert va.lib.vg_query_set_outputs(q.h, ptrs, 5000) == va.VIRG_FAIL
        va.lib.vg_query_free(q.h)          # refused: q stays usable
        b.close()
    pb = va.PersistentBatch([q, q2], direct=1, db=gpu_db)
    pb.reload([q2])                        # lets go of q
    assert va.lib.vg_query_set_outputs(q.h, ptrs, 5000) == va.VIRG_SUCCESS
    assert va.lib.vg_query_set_outputs(q2.h, ptrs, 5000) == va.VIRG_FAIL
    pb.close()
    n = q.run()
    want, _ = oracle.select(gpu_db, "select id from test where uniformi < 50")
    assert n == len(want["id"])
    q.close()
    q2.close()
    gpu_db.device_free(buf)


def test_jit_batch(mix):
    """One persistent launch of the batch's programs compiled together:
    every query's result in row order, bit for bit the oracle's, on every
    graph replay, with eager launches of the same queries in between; the
    batch keeps its own control blocks, so the eager runs do not disturb
    its outcome."""
    qs, want = mix
    t0 = time.perf_counter()
    b = va.PersistentBatch(qs, jit=1)
    print(f"jit batch built in {time.perf_counter() - t0:.1f}s", flush=True)
    try:
        assert b.config()["jit"] == 1
        for replay in range(3):
            print(f"jit batch replay {replay}", flush=True)
            t0 = time.perf_counter()
            rows, lat, flags = b.results()
            wall_ms = (time.perf_counter() - t0) * 1e3
            assert not any(flags), flags
            _check(qs, want, rows, f"jit replay {replay}")
            assert all(0 < x <= wall_ms for q, x in zip(qs, lat) if q.tiles), (wall_ms, lat)
            for i in range(replay, len(qs), 11):
                qs[i].run()
                assert_same(qs[i].fetch(), want[i], f"query {i} eager")
        ms = b.time(1, 3)
        assert ms > 0
        rows, _, flags = b.results()
        assert not any(flags)
        _check(qs, want, rows, "jit after timing")
    finally:
        b.close()


def test_jit_batch_direct_and_reload(mix):
    """Direct (uncaptured) runs, a reload to a reordered subset (a module
    of fewer programs), and a switch back to the interpreter's batch."""
    qs, want = mix
    b = va.PersistentBatch(qs[:20], jit=1, direct=1)
    try:
        rows, _, flags = b.results()
        assert not any(flags)
        _check(qs[:20], want[:20], rows, "jit direct")
        sub = [qs[i] for i in (7, 3, 41, 0, 63, 12)]
        sw = [want[i] for i in (7, 3, 41, 0, 63, 12)]
        b.reload(sub)
        for replay in range(2):
            rows, _, flags = b.results()
            assert not any(flags)
            _check(sub, sw, rows, f"jit reload replay {replay}")
        b.configure(tiles_per_item=4)
        rows, _, flags = b.results()
        assert not any(flags)
        _check(sub, sw, rows, "interpreter after jit")
        b.configure(jit=1)
        rows, _, flags = b.results()
        _check(sub, sw, rows, "jit again")
    finally:
        b.close()


def test_jit_batch_capacity(gpu_db):
    """As the single pass: a result that does not fit is reported with its
    exact row count and never written past the capacity; one that fits
    exactly is whole; an empty result is 0 rows."""
    gpu_db.gen_table("test", 300_001, 1 << 16, seed=7)
    sqls = ["select id, uniformf from test where uniformi < 30",
            "select normalf5 from test where uniformf < 0.5",
            "select id from test where normali20 > 100"]
    exact = []
    for sql in sqls:
        w, _ = oracle.select(gpu_db, sql)
        exact.append((w, len(next(iter(w.values())))))
    assert exact[2][1] == 0
    qs = [gpu_db.prepare(sqls[0], capacity=exact[0][1]),
          gpu_db.prepare(sqls[1], capacity=max(exact[1][1] - 1, 1)),
          gpu_db.prepare(sqls[0], capacity=1000),
          gpu_db.prepare(sqls[2], capacity=1)]
    b = va.PersistentBatch(qs, jit=1)
    try:
        for replay in range(2):
            rows, _, flags = b.results()
            assert [f & 1 for f in flags] == [0, 1, 1, 0]
            assert rows[:3] == [exact[0][1], exact[1][1], exact[0][1]]
            assert rows[3] == 0
            assert_same(qs[0].fetch(), exact[0][0], f"exact fit replay {replay}")
    finally:
        b.close()
        for q in qs:
            q.close()
