"""Batched execution: many prepared queries of different programs over
tables of different sizes, run together, give each query exactly the
result it gives alone (and the oracle's, bit for bit and in row order),
on every replay, with eager launches of the same queries in between."""
from __future__ import annotations

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
