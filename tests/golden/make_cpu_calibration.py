"""Write tests/golden/cpu_calibration.json: the oracle's CPU VM (the
restatement bench.py times as its CPU baseline) against the reference's
own CPU VM (oracle/_ref, built from /root/reference by oracle/build_ref.sh)
on identical inputs, in the development container, so the baseline
measured on the GPU host can be traced to the reference.

Inputs: the 7-column benchmark table from the seeded generator, ROWS rows
in the reference's 8 MB tablets (the reference inserts row by row through
virg_table_insert; the restatement reads the same tablet layout).  Query:
the config-2 math projection at T = 10 and config 1, full results
materialised.  Both VMs run with `threads` pthreads taking whole tablets
(src/vm/vm_cpu.c:58-91) and single-threaded, median of REPS runs, wall
clock around the execute call (as src/test/timing.cc:58-64).

    python tests/golden/make_cpu_calibration.py
"""
from __future__ import annotations

import json
import os
import platform
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import virginian_amd as va                                   # noqa: E402
from oracle import oracle, ref                               # noqa: E402

OUT = Path(__file__).resolve().parent / "cpu_calibration.json"
ROWS = 4 << 20
REPS = 10
COLS = ("uniformi", "normali5", "normali20", "uniformf", "normalf5", "normalf20")
QUERIES = {
    "config2_T10": ("select (uniformf * 2.0) + normalf5 as a, uniformi - normali20 as b, "
                    "uniformf / (normalf20 + 1000.0) as c from test where uniformf < 10.0"),
    "config1": "select id, uniformi, normalf5 from test where uniformi < 50 and normali5 > 0",
}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def median_ms(fn) -> float:
    fn()
    ts = []
    for _ in range(REPS):
        t0 = time.perf_counter()
        fn()
        ts.append((time.perf_counter() - t0) * 1e3)
    return statistics.median(ts)


def main() -> int:
    if not ref.available():
        print("oracle/_ref/libvirgref.so is not built (oracle/build_ref.sh)", file=sys.stderr)
        return 1
    threads = len(os.sched_getaffinity(0))
    c = {k: np.empty(ROWS, t) for k, t in (("id", np.int64), ("uniformi", np.int32), ("normali5", np.int32),
                                            ("normali20", np.int32), ("uniformf", np.float32),
                                            ("normalf5", np.float32), ("normalf20", np.float32))}
    va.lib.virg_gen_rows_mt(1, 0, ROWS, *[x.ctypes.data for x in c.values()], threads)
    db = va.Virginian()
    db.gen_table("test", ROWS, 0, seed=1, threads=threads)       # 8 MB reference tablets
    out = {"rows": ROWS, "tablets": "reference 8 MB tablets", "threads": threads, "cpu": cpu_model(),
           "reps": REPS, "statistic": "median ms, full results materialised", "queries": {}}
    for mode, nthreads in (("multi", threads), ("single", 0)):
        r = ref.RefDB(multi=mode == "multi", threads=max(nthreads, 1))
        try:
            r.table("test", c["id"], [(k, c[k]) for k in COLS])
            for name, sql in QUERIES.items():
                vm = db.compile(sql)
                arr, n = va.ops_array(vm)
                names, types = va.result_schema(vm)
                db.free_vm(vm)
                ref_out = r.run((arr, n), names, types, ROWS)
                ours, _ = oracle.select(db, sql, threads=nthreads, ordered=nthreads == 0)
                if nthreads == 0:
                    assert all(np.array_equal(ref_out[k].view(np.uint8), ours[k].view(np.uint8)) for k in ours)
                t_ref = median_ms(lambda: r.run((arr, n), names, types, ROWS))
                t_ours = median_ms(lambda: oracle.select(db, sql, threads=nthreads, ordered=nthreads == 0))
                out["queries"].setdefault(name, {})[mode] = {
                    "reference_ms": round(t_ref, 2), "restatement_ms": round(t_ours, 2),
                    "restatement_over_reference": round(t_ours / t_ref, 3)}
                print(name, mode, out["queries"][name][mode], flush=True)
        finally:
            r.close()
    db.close()
    out["command"] = "python tests/golden/make_cpu_calibration.py"
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
