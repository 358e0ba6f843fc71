"""A wide parity sweep: 1500 seeded random op-programs (up to 40 ops:
forward jumps, OR groups, keep and drop branches, masked redefinitions,
Invalid, math in every type) over five tables of different sizes and
tablet layouts (ragged last tablets, tablets smaller than a tile), each
run as a resident query and compared bit for bit with the oracle; and 256
of them as persistent batches of 64."""
from __future__ import annotations

import numpy as np
import pytest

import virginian_amd as va
from oracle import oracle
from tests.helpers import assert_same
from virginian_amd import workload as W

pytestmark = pytest.mark.gpu

LAYOUTS = [(3000, 1000), (50_001, 4096), (300_000, 1 << 16), (1_100_000, 1 << 20), (777, 64)]


@pytest.fixture(scope="module")
def sweep():
    db = va.Virginian()
    if db.device()[1] == 0:
        db.close()
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X box")
    tids = [db.gen_table(f"t{k}", n, tr, seed=40 + k) for k, (n, tr) in enumerate(LAYOUTS)]
    rng = np.random.default_rng(2024)
    progs = []
    for i in range(1500):
        sel = float(rng.uniform(0.0001, 0.9)) if i % 4 else None
        progs.append(W.random_program(rng, max_ops=40, selectivity=sel, table=tids[i % len(tids)]))
    yield db, progs
    db.close()


def test_random_programs_resident(sweep):
    db, progs = sweep
    checked = 0
    for i, ops in enumerate(progs):
        want, ill = oracle.select(db, ops)
        if ill:
            continue                # reads a register it never wrote: no defined result
        q = db.prepare(ops)
        try:
            q.run()
            assert_same(q.fetch(), want, f"program {i}")
        finally:
            q.close()
        checked += 1
        if i % 500 == 0:
            print(f"sweep: {i} programs", flush=True)
    assert checked > 1200


def test_random_programs_batched(sweep):
    db, progs = sweep
    for s in range(0, 256, 64):
        chunk = progs[s:s + 64]
        print(f"sweep batch at {s}", flush=True)
        qs = [db.prepare(p) for p in chunk]
        b = va.PersistentBatch(qs, tiles_per_item=8)
        try:
            b.run()
            for j, (q, ops) in enumerate(zip(qs, chunk)):
                want, ill = oracle.select(db, ops)
                if not ill:
                    assert_same(q.fetch(), want, f"batched program {s + j}")
        finally:
            b.close()
            for q in qs:
                q.close()
