"""GPU parity on seeded random op programs (forward jumps, OR groups with
waiting rows, masked redefinitions, Invalid, Rowid, float division) and
on hand-built programs for the VM's less common ops and types, against
the CPU oracle, bit for bit and in row order."""
from __future__ import annotations

import numpy as np
import pytest

import virginian_amd as va
from oracle import oracle
from tests.helpers import assert_same, cast_programs, typed_program, typed_table
from virginian_amd import workload as W

pytestmark = pytest.mark.gpu


def _gpu(db, prog):
    q = db.prepare(prog)
    try:
        q.run()
        return q.fetch()
    finally:
        q.close()


@pytest.mark.parametrize("tablet_rows", [0, 1 << 16])
def test_random_programs(gpu_db, tablet_rows):
    """120 programs of up to 40 ops over 300k rows (reference 8 MB tablets,
    and 64k-row tablets with a ragged last one)."""
    gpu_db.gen_table("test", 300_001, tablet_rows, seed=11)
    rng = np.random.default_rng(77 + tablet_rows)
    for i in range(120):
        sel = float(rng.uniform(0.0001, 0.8)) if rng.random() < 0.7 else None
        prog = W.random_program(rng, max_ops=40, selectivity=sel)
        print(f"program {i}: {len(prog)} ops", flush=True)
        want, ill = oracle.select(gpu_db, prog)
        assert ill == 0
        got = _gpu(gpu_db, prog)
        assert_same(got, want, f"program {i}: {prog}")


def test_types_and_ops(gpu_db):
    """char / double / int64 arithmetic (char results wrap to 8 bits), And,
    Or, Not, Neq, masked redefinition by Integer, integer division, Cast
    (the oracle's result is itself pinned against numpy on the CPU)."""
    typed_table(gpu_db, 5000, 3)
    ops = typed_program()
    want, ill = oracle.select(gpu_db, ops)
    assert ill == 0
    assert len(want["c0"]) > 1000
    got = _gpu(gpu_db, ops)
    assert_same(got, want, "typed program")


def test_casts(gpu_db):
    """The Cast programs whose oracle results equal the reference VM's own
    (tests/test_ref_parity.py): same-width and narrowing conversions
    between every numeric type, rows dropped before the Casts."""
    typed_table(gpu_db, 70_000, 21)
    for k, ops in enumerate(cast_programs()):
        want, ill = oracle.select(gpu_db, ops)
        assert ill == 0
        got = _gpu(gpu_db, ops)
        assert_same(got, want, f"cast program {k}")


def test_capacity_overflow_is_an_error(gpu_db):
    """A result larger than the caller's capacity fails the query instead
    of writing past the output buffers."""
    gpu_db.gen_table("test", 100_000, 1 << 15, seed=4)
    q = gpu_db.prepare("select id from test where uniformi < 50", capacity=1000)
    try:
        with pytest.raises(va.VirginianError):
            q.run()
        # the query stays usable after the failure
        with pytest.raises(va.VirginianError):
            q.run()
    finally:
        q.close()


def test_type_mismatch_rejected(gpu_db):
    """int against float operands abort the reference (virginia.c:52); the
    executor refuses the program up front."""
    gpu_db.gen_table("test", 1000, 0, seed=1)
    with pytest.raises(va.VirginianError):
        gpu_db.prepare("select id from test where uniformi < 2.0")
