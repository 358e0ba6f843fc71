"""The serving loop: requests arriving faster than they run are batched
while others are on the GPU, admitted by the HBM their buffers take, and
each result is the same as when the query runs alone (the oracle's, bit
for bit and in row order), in single-pass mode with reruns of results
that did not fit and in two-pass mode."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from tests.helpers import assert_same
from virginian_amd import predict as PR
from virginian_amd import serve as S
from virginian_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture
def setup(gpu_db):
    rng = np.random.default_rng(31)
    tids = []
    for k, rows in enumerate([150_000, 600_000, 2_000_000]):
        gpu_db.gen_table(f"s{k}", rows, 1 << 18, seed=60 + k)
        tids.append(gpu_db.table_id(f"s{k}"))
    reqs, t = [], 0.0
    for i in range(72):
        t += float(rng.exponential(1 / 20000.0))      # arrivals faster than service
        sel = float(rng.uniform(0.0001, 0.8))
        ops = W.random_program(rng, max_ops=40, selectivity=sel, table=int(rng.choice(tids)))
        reqs.append(S.Request(arrival=t, ops=ops, rid=i))
    want = [oracle.select(gpu_db, r.ops)[0] for r in reqs]
    return gpu_db, tids, reqs, want


def _fresh(reqs):
    return [S.Request(arrival=r.arrival, ops=r.ops, rid=r.rid) for r in reqs]


def _check(done, want, what):
    for r, w in zip(sorted(done, key=lambda r: r.rid), want):
        assert r.rows == len(next(iter(w.values()))), f"request {r.rid} {what}"
        assert_same(r.result, w, f"request {r.rid} {what}")


@pytest.mark.parametrize("mode,margin", [("single", 0.0), ("single", 0.2), ("sized", 0.0), ("two-pass", 0.0),
                                         ("plain", 0.0)])
def test_served_results_equal_alone(setup, mode, margin):
    db, tids, reqs, want = setup
    srv = S.Server(db, tids, 256 << 20, mode=mode, predictor=None, margin=margin, runners=3, max_batch=16)
    try:
        done = _fresh(reqs)
        summ = srv.run(done)
    finally:
        srv.close()
    print(mode, margin, summ, flush=True)
    assert all(r.ready >= r.arrival for r in done)
    _check(done, want, f"{mode} margin {margin}")
    if mode in ("single", "sized") and margin == 0.0:
        assert summ["reruns"] > 0                 # some estimates were short: reruns exercised


@pytest.mark.parametrize("mode", ["two-pass", "sized", "single"])
def test_served_results_into_host_memory(setup, mode):
    """With the arena in pinned host memory the kernels write each result
    straight to the host; the results are the same as alone."""
    db, tids, reqs, want = setup
    srv = S.Server(db, tids, 256 << 20, mode=mode, predictor=None, margin=0.05, runners=3, max_batch=16,
                   host_arena=True)
    try:
        done = _fresh(reqs)
        summ = srv.run(done)
    finally:
        srv.close()
    print(mode, "host arena", summ, flush=True)
    _check(done, want, f"{mode} host arena")


def test_alone_equals_in_mix(setup):
    """One at a time (one runner, batches of one) gives the same results
    as the mix."""
    db, tids, reqs, want = setup
    srv = S.Server(db, tids, 256 << 20, mode="single", margin=0.1, runners=1, max_batch=1)
    try:
        done = _fresh(reqs[:24])
        srv.run(done)
    finally:
        srv.close()
    _check(done, want[:24], "alone")


def test_admission_by_footprint(setup):
    """With a small arena, queries wait for memory; the arena is never
    over-committed and every request is served."""
    db, tids, reqs, want = setup
    arena = 48 << 20
    srv = S.Server(db, tids, arena, mode="two-pass", runners=3, max_batch=16)
    try:
        done = _fresh(reqs[:30])
        summ = srv.run(done)
    finally:
        srv.close()
    assert summ["peak_arena_bytes"] <= arena
    _check(done, want[:30], "small arena")


def test_served_with_trained_predictor(setup):
    db, tids, reqs, want = setup
    stats = {t: PR.TableStats(db, t) for t in tids}
    rng = np.random.default_rng(9)
    train = [W.random_program(rng, max_ops=40, selectivity=float(rng.uniform(0.0001, 0.8)),
                              table=int(rng.choice(tids))) for _ in range(120)]
    x, t = PR.collect(db, train[:-16], stats)
    xv, tv = PR.collect(db, train[-16:], stats)
    net = PR.MLP(seed=2)
    net.fit(x, t, xv, tv)
    srv = S.Server(db, tids, 256 << 20, mode="single", predictor=net, margin=0.05, runners=3, max_batch=16)
    try:
        done = _fresh(reqs)
        summ = srv.run(done)
    finally:
        srv.close()
        net.close()
    print(summ, flush=True)
    _check(done, want, "predictor-sized")
