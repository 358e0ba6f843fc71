"""Shared test helpers: reference fixtures and exact comparison."""
from __future__ import annotations

import numpy as np

import virginian_amd as va

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
QUERY_SUITE = [
    "select id, uniformi, normali5 from test where uniformi > 60 and normali5 < 0",
    "select id, uniformf, normalf5 from test where uniformf > 60.0 and normalf5 < 0.0",
    "select id, uniformi, normali5 from test where uniformi > -60 and normali5 < 5",
    "select id, uniformf, normalf5 from test where uniformf > -60.0 and normalf5 < 5.0",
    "select id, uniformi, normali20 from test where (normali20 + 40) > (uniformi - 10)",
    "select id, uniformf, normalf20 from test where (normalf20 + 40.0) > (uniformf - 10.0)",
    "select id, normali5, normali20 from test where normali5 * normali20 >= -5 and normali5 * normali20 <= 5",
    "select id, normalf5, normalf20 from test where normalf5 * normalf20 >= -5.0 and normalf5 * normalf20 <= 5.0",
    "select id, uniformi, normali5, normali20 from test where uniformi >= -1 and uniformi <= 1 or normali5 >= -1 and normali5 <= 1 or normali20 >= -1 and normali20 <= 1",
    "select id, uniformf, normalf5, normalf20 from test where uniformf >= -1.0 and uniformf <= 1.0 or normalf5 >= -1.0 and normalf5 <= 1.0 or normalf20 >= -1.0 and normalf20 <= 1.0",
]

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
