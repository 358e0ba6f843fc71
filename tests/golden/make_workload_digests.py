"""Write tests/golden/workload_digests.json: what the benchmark workloads
must produce, computed by the CPU oracle in the development container.

For each workload: rows selected, the result digest (virginian_amd/verify.py)
and the rows selected in every tablet.
  config 1   1,048,576 rows, one tablet;
  config 2   67,108,864 rows in 1M-row tablets, thresholds 0.1/10/50/90 %;
  config2_1M the same queries over the first 1,048,576 rows (GPU tests);
  config 3   2,000,000,000 rows in 1M-row tablets (1908, the last partial),
             generated and run a few tablets at a time.
Tables come from the seeded generator (seed 1, as bench.py).  At 1M rows
the oracle's results are checked against the reference's own CPU VM
(oracle/_ref) where it is built; the check's outcome is recorded.

    python tests/golden/make_workload_digests.py [--skip-config3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import virginian_amd as va                                   # noqa: E402
from oracle import oracle, ref                               # noqa: E402
from virginian_amd.verify import Digest                      # noqa: E402

OUT = Path(__file__).resolve().parent / "workload_digests.json"
SEED = 1
M = 1 << 20
CONFIG1_SQL = "select id, uniformi, normalf5 from test where uniformi < 50 and normali5 > 0"
CONFIG3_SQL = ("select id, uniformi, normalf20 from test where uniformf < 9.24 and normalf5 > 2.0 "
               "or normali20 > 10")
THRESHOLDS = [0.1, 10.0, 50.0, 90.0]
COLS = ("uniformi", "normali5", "normali20", "uniformf", "normalf5", "normalf20")
THREADS = max(1, len(os.sched_getaffinity(0)))


def config2_sql(t: float) -> str:
    return ("select (uniformf * 2.0) + normalf5 as a, uniformi - normali20 as b, "
            f"uniformf / (normalf20 + 1000.0) as c from test where uniformf < {t!r}")


def run(db, sql: str, ntab: int) -> dict:
    """Digest, rows and per-tablet rows of sql over db's table "test"."""
    d, per = None, []
    for i in range(ntab):
        got, ill = oracle.select(db, sql, threads=THREADS, tablets=(i, i + 1))
        assert ill == 0
        cols = list(got.values())
        if d is None:
            d = Digest(len(cols))
        d.update(cols)
        per.append(len(cols[0]))
    return {"selected": d.rows, "digest": d.hexdigest(), "per_tablet": per}


def gen_rows(row0: int, n: int):
    cols = {"id": np.empty(n, np.int64), "uniformi": np.empty(n, np.int32),
            "normali5": np.empty(n, np.int32), "normali20": np.empty(n, np.int32),
            "uniformf": np.empty(n, np.float32), "normalf5": np.empty(n, np.float32),
            "normalf20": np.empty(n, np.float32)}
    va.lib.virg_gen_rows_mt(SEED, row0, n, *[c.ctypes.data for c in cols.values()], THREADS)
    return cols


def reference_check(db, sqls: list[str], rows: int) -> list[bool]:
    """The reference VM over the same rows: results equal the oracle's?"""
    if not ref.available():
        return []
    c = gen_rows(0, rows)
    r = ref.RefDB()
    ok = []
    try:
        r.table("test", c["id"], [(k, c[k]) for k in COLS])
        for sql in sqls:
            vm = db.compile(sql)
            arr, n = va.ops_array(vm)
            names, types = va.result_schema(vm)
            db.free_vm(vm)
            got = r.run((arr, n), names, types, rows)
            want, _ = oracle.select(db, sql, threads=THREADS)
            ok.append(all(np.array_equal(got[k].view(np.uint8), want[k].view(np.uint8)) for k in want))
    finally:
        r.close()
    return ok


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-config3", action="store_true")
    a = ap.parse_args()
    doc = json.loads(OUT.read_text()) if OUT.exists() else {}
    doc["digest"] = ("sha256 over the concatenated sha256 digests of the result columns, each over the "
                     "column's bytes in result order (virginian_amd/verify.py)")
    doc["seed"] = SEED
    t0 = time.time()

    db = va.Virginian()
    db.gen_table("test", M, M, seed=SEED, threads=THREADS)
    doc["config1"] = {"rows": M, "tablet_rows": M, "sql": CONFIG1_SQL, **run(db, CONFIG1_SQL, 1)}
    doc["config2_1M"] = {"rows": M, "tablet_rows": M,
                         "queries": [{"threshold": t, **run(db, config2_sql(t), 1)} for t in THRESHOLDS]}
    doc["reference_check_1M"] = {"queries": [CONFIG1_SQL] + [config2_sql(t) for t in THRESHOLDS],
                                 "equal": reference_check(db, [CONFIG1_SQL] + [config2_sql(t) for t in THRESHOLDS], M)}
    db.close()
    print(f"config 1 and 1M done ({time.time() - t0:.0f}s); reference check {doc['reference_check_1M']['equal']}",
          flush=True)

    db = va.Virginian()
    db.gen_table("test", 64 * M, M, seed=SEED, threads=THREADS)
    doc["config2"] = {"rows": 64 * M, "tablet_rows": M,
                      "queries": [{"threshold": t, **run(db, config2_sql(t), 64)} for t in THRESHOLDS]}
    db.close()
    print(f"config 2 done ({time.time() - t0:.0f}s)", flush=True)
    OUT.write_text(json.dumps(doc, indent=1) + "\n")

    if not a.skip_config3:
        rows = 2_000_000_000
        ntab = (rows + M - 1) // M
        d, per = Digest(3), []
        step = 16
        for t in range(0, ntab, step):
            db = va.Virginian()
            tid = db.create_table("test", va.VIRG_INT64, [(k, va.VIRG_INT if k in COLS[:3] else va.VIRG_FLOAT)
                                                          for k in COLS])
            nt = min(step, ntab - t)
            for k in range(t, t + nt):
                n = min(M, rows - k * M)
                c = gen_rows(k * M, n)
                db.append_tablet(tid, c["id"], [c[x] for x in COLS])
            for i in range(nt):
                got, ill = oracle.select(db, CONFIG3_SQL, threads=THREADS, tablets=(i, i + 1))
                assert ill == 0
                cols = list(got.values())
                d.update(cols)
                per.append(len(cols[0]))
            db.close()
            if t % 256 == 0:
                print(f"config 3: tablet {t}/{ntab} ({time.time() - t0:.0f}s)", flush=True)
        doc["config3"] = {"rows": rows, "tablet_rows": M, "sql": CONFIG3_SQL,
                          "selected": d.rows, "digest": d.hexdigest(), "per_tablet": per}
        OUT.write_text(json.dumps(doc, indent=1) + "\n")
        print(f"config 3 done ({time.time() - t0:.0f}s)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
