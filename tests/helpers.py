"""Shared test helpers: reference fixtures and exact comparison."""
from __future__ import annotations

import numpy as np

import virginian_amd as va
from virginian_amd.workload import SUITE_QUERIES

# reference: src/test/sql.cc:15-37
KAT_CONDITIONS = [
    "select col0 from test where col0 < 9 and col0 >= 7 or col0 = 3",
    "select col0 from test where col0 >= 8 or col0 = 3 and col0 < 1",
    "select col0 from test where col0 = 8 and col0 = 8",
    "select col0 from test where col0 = 8 or col0 = 8",
    "select col0 from test where col0 = 2",
    "select col0 from test where (col0 = 2)",
    "select col0 from test where col0 < 5 and (col0 = 7 or col0 = 2)",
    "select col0 from test where col0 = 7 or col0 = 2 or col0 = 8",
    "select col0 from test where col0 < 5 and (col0 = 2 or col0 = 8 or col0 = 1)",
]
KAT_CONDITION_ROWS = [3, 2, 1, 1, 1, 1, 1, 3, 2]

# reference: src/test/queries.h:1-13 (the published 10-query suite)
QUERY_SUITE = SUITE_QUERIES

# the benchmark workloads
CONFIG1_SQL = "select id, uniformi, normalf5 from test where uniformi < 50 and normali5 > 0"
CONFIG2_THRESHOLDS = [0.1, 10.0, 50.0, 90.0]


def config2_sql(threshold: float, table: str = "test") -> str:
    return ("select (uniformf * 2.0) + normalf5 as a, uniformi - normali20 as b, "
            f"uniformf / (normalf20 + 1000.0) as c from {table} where uniformf < {threshold!r}")


def simpledb(db: va.Virginian, rows: int) -> int:
    """The reference test fixture: table "test", int key, col0..col2 =
    i, i+1, i+2 (reference: src/test/test.cc:12-40)."""
    tid = db.create_table("test", va.VIRG_INT, [("col0", va.VIRG_INT), ("col1", va.VIRG_INT),
                                                  ("col2", va.VIRG_INT)])
    i = np.arange(rows, dtype=np.int32)
    db.insert(tid, i, [i, i + 1, i + 2])
    return tid


def assert_same(got: dict, want: dict, what: str = "") -> None:
    """Bit-exact equality of result columns (floats compared as bits)."""
    assert list(got) == list(want), (what, list(got), list(want))
    for k in want:
        g, w = got[k], want[k]
        assert g.dtype == w.dtype, (what, k, g.dtype, w.dtype)
        assert len(g) == len(w), f"{what} column {k}: {len(g)} rows, oracle {len(w)}"
        if not np.array_equal(g.view(np.uint8), w.view(np.uint8)):
            bad = np.nonzero(g.view(f"u{g.dtype.itemsize}") != w.view(f"u{w.dtype.itemsize}"))[0]
            raise AssertionError(f"{what} column {k}: {len(bad)} mismatches, first at {bad[:5]}: "
                                 f"got {g[bad[:5]]} want {w[bad[:5]]}")


def op_program(result_types, body, nres_reg):
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


def typed_table(db, rows, seed):
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


def cast_programs() -> list[list[tuple]]:
    """Programs over typed_table() whose Casts the reference runs the same
    way as the clean definition: the reference's Cast takes its type and
    register from the last op that fetched operands (virginia.c:682-862),
    so each Cast here follows an Integer with the same p1 and p2
    (Integer r<type> = <register>, a harmless write to a low register);
    only conversions that change the type and do not widen the value
    (the reference converts in place in row order, so a 4-to-8-byte
    widening overwrites rows before it reads them, and a cast to the same
    type leaves its rows behind, which the reference's own assertion at
    virginia.c:376 rejects); no double or float is cast to char (out of
    range there is undefined in C).  The second program first drops the
    rows with i > 300, which then skip the Casts."""
    O = va

    def cast(to, reg, col):
        return [(O.OP_Column, reg, col, 0, None), (O.OP_Integer, to, reg, 0, None),
                (O.OP_Cast, to, reg, 0, None)]

    # typed_table columns: 0 c (char), 1 d (double), 2 l (int64), 3 i (int), 4 f (float)
    first = (cast(O.VIRG_INT, 10, 1) + cast(O.VIRG_INT, 11, 4) + cast(O.VIRG_FLOAT, 12, 3)
             + cast(O.VIRG_DOUBLE, 13, 2) + cast(O.VIRG_INT64, 14, 1) + cast(O.VIRG_FLOAT, 15, 2))
    second = ([(O.OP_Column, 9, 3, 0, None), (O.OP_Integer, 8, 300, 0, None), (O.OP_Gt, 9, 8, "R", 0)]
              + cast(O.VIRG_INT, 10, 2) + cast(O.VIRG_FLOAT, 11, 1) + cast(O.VIRG_CHAR, 12, 3)
              + cast(O.VIRG_CHAR, 13, 2) + cast(O.VIRG_INT64, 14, 1) + cast(O.VIRG_INT, 15, 4))
    return [op_program([O.VIRG_INT, O.VIRG_INT, O.VIRG_FLOAT, O.VIRG_DOUBLE, O.VIRG_INT64,
                        O.VIRG_FLOAT], first, 10),
            op_program([O.VIRG_INT, O.VIRG_FLOAT, O.VIRG_CHAR, O.VIRG_CHAR, O.VIRG_INT64,
                        O.VIRG_INT], second, 10)]


def typed_program() -> list[tuple]:
    """A program over typed_table() exercising char, double and int64
    arithmetic, Cast, And/Or/Not/Neq branches that keep validity, a
    masked redefinition by Integer and integer division; results r5..r12."""
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
    return op_program([O.VIRG_CHAR, O.VIRG_DOUBLE, O.VIRG_INT64, O.VIRG_INT, O.VIRG_INT,
                       O.VIRG_DOUBLE, O.VIRG_DOUBLE, O.VIRG_INT64], body, 5)
