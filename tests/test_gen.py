"""The synthetic table generator (host C): known-answer vectors of its
counter-based generator, accuracy of its logarithm, the distributions the
benchmark workloads are quoted on, and determinism."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import virginian_amd as va
from tests.test_oracle import gen_columns

# Philox4x32-10 known answers published with Random123 (Salmon et al., SC'11,
# kat_vectors): (counter, key) -> output
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_kat(ctr, key, want):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    va.lib.virg_philox(c, k, o)
    assert tuple(o) == want


def test_log_accuracy():
    """virg_log is written from add/multiply/divide only; it must agree
    with the C library's log to within 2 ulp over (0, 1]."""
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.random(20000), 10.0 ** rng.uniform(-300, 0, 20000),
                         [1.0, 0.5, 0.25, 1e-300, 0.7071067811865476, 0.7071067811865475,
                          np.nextafter(1.0, 0.0)]])
    worst = 0.0
    for x in xs:
        x = float(x)
        if x <= 0.0:
            continue
        got, want = va.lib.virg_log(x), math.log(x)
        if want == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, abs(got - want) / math.ulp(want))
    assert worst <= 2.0, worst


def test_distributions():
    c = gen_columns(1 << 20, 11)
    n = len(c["id"])
    assert np.array_equal(c["id"], np.arange(n))
    ui = c["uniformi"]
    assert ui.min() == 0 and ui.max() == 99
    hist = np.bincount(ui, minlength=100)
    chi2 = float(((hist - n / 100) ** 2 / (n / 100)).sum())
    assert chi2 < 160, chi2           # 99 dof: p ~ 1e-4 at 160
    uf = c["uniformf"]
    assert uf.min() >= 0.0 and uf.max() < 100.0
    assert abs(float(uf.mean()) - 50.0) < 0.2
    for name, sigma in (("normalf5", 5.0), ("normalf20", 20.0)):
        x = c[name].astype(np.float64)
        assert abs(x.mean()) < 0.02 * sigma
        assert abs(x.std() / sigma - 1.0) < 0.005
        # tails of a normal: P(|N| > 3) = 0.0027
        assert abs(float((np.abs(x) > 3 * sigma).mean()) - 0.0027) < 0.0004
    for name, sigma in (("normali5", 5.0), ("normali20", 20.0)):
        x = c[name]
        assert abs(float(x.mean())) < 0.02 * sigma
    # the config-1 selectivity: P(uniformi < 50) * P(round(5N) > 0) = 0.5 * P(N >= 0.1)
    p = 0.5 * 0.5 * math.erfc(0.1 / math.sqrt(2.0))
    sel = float(((ui < 50) & (c["normali5"] > 0)).mean())
    assert abs(sel - p) < 0.002, (sel, p)


def test_rounding_half_away_from_zero():
    """normali = round(N * sigma) half away from zero: the integer columns
    are within 0.5 of sigma times some normal and symmetric about 0."""
    c = gen_columns(1 << 18, 12)
    for name in ("normali5", "normali20"):
        x = c[name]
        assert abs(int((x > 0).sum()) - int((x < 0).sum())) < 4 * math.sqrt(len(x))


def test_deterministic_and_position_independent():
    a = gen_columns(50_000, 9)
    b = {k: np.empty(20_000, v.dtype) for k, v in a.items()}
    va.lib.virg_gen_rows_mt(9, 30_000, 20_000, *[v.ctypes.data for v in b.values()], 3)
    for k in a:
        assert np.array_equal(a[k][30_000:].view(np.uint8), b[k].view(np.uint8)), k
    c = gen_columns(50_000, 10)
    assert not np.array_equal(a["uniformf"], c["uniformf"])


def test_device_builder_needs_a_device():
    """vg_gen_table has no host fallback: without a device it fails."""
    db = va.Virginian()
    try:
        if db.device()[1] > 0:
            pytest.skip("a device is present (the device builder is tested in test_gpu_gen.py)")
        with pytest.raises(va.VirginianError):
            db.gen_table("t", 100, 100, seed=1, device=True)
    finally:
        db.close()


def test_reference_suite_table():
    """The reference's test table (db/generate.c:41-48): int key 0..n-1,
    C-truncated ints of U(-100, 100) and N(0, sigma), f32 floats of the
    same.  The suite's first query selects trunc(U) > 60 and trunc(N(0, 5))
    < 0, with probability 39/200 x Phi(-1/5) (trunc(x) < 0 iff x <= -1 for
    the integer draws, trunc(u) > 60 iff u >= 61), checked through the
    oracle against that closed form."""
    from oracle import oracle
    from virginian_amd.workload import SUITE_QUERIES, suite_table
    rows = 400_000
    db = va.Virginian()
    try:
        tid = suite_table(db, rows, 1 << 16, seed=3)
        assert db.numrows(tid) == rows
        got, ill = oracle.select(db, "select id, uniformi, normali5, uniformf from test where id >= 0", threads=8)
        assert ill == 0
        assert np.array_equal(got["id"], np.arange(rows, dtype=np.int32))
        u, g5, uf = got["uniformi"], got["normali5"], got["uniformf"]
        assert u.dtype == np.int32 and uf.dtype == np.float32
        assert u.min() == -99 and u.max() == 99               # truncation toward zero, never +-100
        assert abs(np.mean(u == 0) - 2 / 200) < 0.002         # 0 takes (-1, 1): twice a bin's share
        assert abs(np.mean(g5 == 0) - (math.erf(1 / (5 * math.sqrt(2))))) < 0.004
        assert -100 <= uf.min() and uf.max() < 100
        want, ill = oracle.select(db, SUITE_QUERIES[0], threads=8)
        assert ill == 0
        p = 39 / 200 * 0.5 * (1 + math.erf(-0.2 / math.sqrt(2)))
        sel = len(want["id"]) / rows
        assert abs(sel - p) < 4 * math.sqrt(p * (1 - p) / rows), (sel, p)
    finally:
        db.close()
