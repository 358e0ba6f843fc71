"""GPU parity on seeded random op programs (forward jumps, OR groups with
waiting rows, masked redefinitions, Invalid, Rowid, float division) and
on hand-built programs for the VM's less common ops and types, against
the CPU oracle, bit for bit and in row order."""
from __future__ import annotations

import numpy as np
import pytest

import virginian_amd as va
from oracle import oracle
from tests.helpers import assert_same
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


def _prog(result_types, body, nres_reg):
    """Table 0, ResultColumns, Parallel, body, Result(nres_reg...), Converge, Finish."""
    ops = [(va.OP_Table, 0, 0, 0, None)]
    ops += [(va.OP_ResultColumn, t, 0, 0, f"c{j}") for j, t in enumerate(result_types)]
    par = len(ops)
    ops.append((va.OP_Parallel, 0, 0, 0, None))
    base = len(ops)
    for op in body:
        op = list(op)
        # jump targets in `body` are relative to the body; 'R' = Result
        if op[0] in (va.OP_Lt, va.OP_Le, va.OP_Gt, va.OP_Ge, va.OP_Eq, va.OP_Neq,
                     va.OP_And, va.OP_Or, va.OP_Not):
            op[3] = base + len(body) if op[3] == "R" else base + op[3]
        ops.append(tuple(op))
    ops.append((va.OP_Result, nres_reg, len(result_types), 0, None))
    ops.append((va.OP_Converge, 0, 0, 0, None))
    ops.append((va.OP_Finish, 0, 0, 0, None))
    ops[par] = (va.OP_Parallel, 0, 0, len(ops) - 1, None)
    return ops


def _typed_table(db, rows, seed):
    """A table with char, double and int64 columns besides int and float."""
    rng = np.random.default_rng(seed)
    tid = db.create_table("test", va.VIRG_INT64, [("c", va.VIRG_CHAR), ("d", va.VIRG_DOUBLE),
                                                  ("l", va.VIRG_INT64), ("i", va.VIRG_INT),
                                                  ("f", va.VIRG_FLOAT)])
    key = np.arange(rows, dtype=np.int64) * 3
    cols = [rng.integers(-128, 128, rows).astype(np.int8),
            rng.normal(0, 100, rows).astype(np.float64),
            rng.integers(-2**40, 2**40, rows).astype(np.int64),
            rng.integers(-1000, 1000, rows).astype(np.int32),
            rng.normal(0, 10, rows).astype(np.float32)]
    db.append_tablet(tid, key, cols)
    return cols


def test_types_and_ops(gpu_db):
    """char / double / int64 arithmetic (char results wrap to 8 bits), And,
    Or, Not, Neq, masked redefinition by Integer, int division including
    by zero and by -1 (defined as 0 and negation here), Cast."""
    _typed_table(gpu_db, 5000, 3)
    O = va
    body = [
        (O.OP_Column, 0, 0, 0, None),          # 0  r0 = c (char)
        (O.OP_Column, 1, 1, 0, None),          # 1  r1 = d (double)
        (O.OP_Column, 2, 2, 0, None),          # 2  r2 = l (int64)
        (O.OP_Column, 3, 3, 0, None),          # 3  r3 = i (int)
        (O.OP_Integer, 4, 7, 0, None),         # 4  r4 = 7
        (O.OP_Gt, 3, 4, "R", 0),               # 5  i > 7: drop the row
        (O.OP_Add, 5, 0, 0, None),             # 6  r5 = c + c (char, wraps)
        (O.OP_Mul, 6, 1, 1, None),             # 7  r6 = d * d
        (O.OP_Sub, 7, 2, 2, None),             # 8  r7 = l - l
        (O.OP_Add, 7, 7, 2, None),             # 9  r7 = r7 + l (redefinition)
        (O.OP_Integer, 8, 3, 0, None),         # 10 r8 = 3
        (O.OP_Neq, 3, 8, 12 + 1, 1),           # 11 i != 3: skip 12, keep validity
        (O.OP_Integer, 8, -1, 0, None),        # 12 rows with i == 3 only: r8 = -1
        (O.OP_Div, 9, 3, 8, None),             # 13 r9 = i / r8
        (O.OP_Column, 10, 4, 0, None),         # 14 r10 = f
        (O.OP_Cast, O.VIRG_DOUBLE, 10, 0, None),   # 15 r10 = (double) f
        (O.OP_Add, 11, 10, 1, None),           # 16 r11 = f + d
        (O.OP_Rowid, 12, 0, 0, None),          # 17 r12 = key
        (O.OP_Integer, 13, 1, 0, None),        # 18 r13 = 1
        (O.OP_Not, 3, 0, 21, 1),               # 19 i == 0: skip 20, keep validity
        (O.OP_Integer, 13, 2, 0, None),        # 20 rows with i != 0: r13 = 2
        (O.OP_Lt, 13, 4, 23, 1),               # 21 r13 < 7 (always): skip the Invalid
        (O.OP_Invalid, 0, 0, 0, None),         # 22
        (O.OP_And, 3, 4, 24, 1),               # 23 i && 7: to 24 either way
        (O.OP_Or, 8, 8, 25, 1),                # 24 r8 || r8: to 25 either way
        (O.OP_Add, 14, 13, 13, None),          # 25 r14 = 2 * r13 (not in the result)
    ]
    # results r5 (char) r6 (double) r7 (int64) r8 r9 (int) r10 r11 (double) r12 (int64)
    ops = _prog([O.VIRG_CHAR, O.VIRG_DOUBLE, O.VIRG_INT64, O.VIRG_INT, O.VIRG_INT,
                 O.VIRG_DOUBLE, O.VIRG_DOUBLE, O.VIRG_INT64], body, 5)
    want, ill = oracle.select(gpu_db, ops)
    assert ill == 0
    assert len(want["c0"]) > 1000
    got = _gpu(gpu_db, ops)
    assert_same(got, want, "typed program")


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
