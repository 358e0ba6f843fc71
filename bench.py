"""Benchmark: filter + math SELECT over a GPU-resident 64M-row table.

Workload: 67,108,864 rows of the 7-column benchmark schema in 1M-row
tablets, resident in HBM;
    select (uniformf * 2.0) + normalf5, uniformi - normali20,
           uniformf / (normalf20 + 1000.0) from test where uniformf < T
for T in 0.1, 10, 50, 90 (selectivity 0.1 %, 10 %, 50 %, 90 %).  One step
runs the four queries, each one launch of its specialised single-pass
kernel over all 64 tablets on its own stream.  value = scanned bytes
(referenced columns at full length) of all queries / wall time of the K
timed steps.  The 64M-row results are checked against the oracle's
digests outside the timed region (verified).

Also reported:
  roofline      the query's kernel: algorithmic bytes per launch
                (referenced columns x rows + selected rows x result bytes)
                / mean launch time from HIP events on the query's stream,
                against 8 TB/s; per query in config.per_query, beside the
                interpreter's count + write launch pair;
  traffic       HBM bytes per launch measured with rocprofv3 FETCH_SIZE /
                WRITE_SIZE, scaled by the same counters on a device copy
                of known size (profiles/traffic_*.json), or null;
  cpu_baseline  the oracle's multi-threaded CPU VM on this host's cores
                over a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --workload config1|outofcore|batched|served|suite
    python bench.py --copy-probe 256     # known-size device copies (counter calibration)
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "GB/s scanned (% HBM peak) for filter SELECT; p50/p99 latency, batched vs plain"
THRESHOLDS = [0.1, 10.0, 50.0, 90.0]


def sql_for(t: float) -> str:
    return ("select (uniformf * 2.0) + normalf5 as a, uniformi - normali20 as b, "
            f"uniformf / (normalf20 + 1000.0) as c from test where uniformf < {t!r}")


def copy_probe(mib: int, reps: int = 5) -> None:
    """Device-to-device copies of `mib` MiB through the HIP runtime: a
    kernel of known traffic (read and write mib MiB) for calibrating the
    FETCH_SIZE / WRITE_SIZE counters."""
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        if hip.hipMemcpy(b, a, C.c_size_t(n), 3) != 0:     # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_probe_mib": mib, "reps": reps, "ms_per_copy": round(dt * 1e3, 3),
                      "read_bytes": n, "write_bytes": n}))
    hip.hipFree(a)
    hip.hipFree(b)


def cpu_threads() -> int:
    """Host threads for the CPU baseline: this process's cores, at most the
    16 a one-GPU host gives a job."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(n, 16))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(db, sql: str, seconds: float, table: str | None = None) -> dict:
    """The oracle's multi-threaded CPU VM (the reference's multi-core
    scheduling: threads take whole tablets, src/vm/vm_cpu.c:58-91) over the
    table, results materialised as the GPU's are, repeated until about
    `seconds` of wall time have run."""
    from oracle import oracle
    threads = cpu_threads()
    reps = 0
    t0 = time.perf_counter()
    while True:
        oracle.select(db, sql, table=table, threads=threads, ordered=False)
        reps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    return {"reps": reps, "seconds": time.perf_counter() - t0, "threads": threads}


def cpu_calibration() -> dict | None:
    """The restatement against the reference's own CPU VM, timed on the same
    inputs in the development container (tests/golden/make_cpu_calibration.py)."""
    p = ROOT / "tests" / "golden" / "cpu_calibration.json"
    return json.loads(p.read_text()) if p.exists() else None


CONFIG1_SQL = "select id, uniformi, normalf5 from test where uniformi < 50 and normali5 > 0"


def bench_config1(a) -> int:
    """The plain filter over one 1M-row tablet (select id, uniformi, normalf5
    where uniformi < 50 and normali5 > 0; 23 % selected): latency, best of
    5 and median of 20, of the resident query's launch (HIP events) and of the
    drop-in virg_vm_execute (wall clock, results into host result tablets),
    beside the oracle's single-core CPU VM on the same table.  At 21 MB
    read the GPU side is launch-bound; the figures are latencies."""
    import statistics
    import virginian_amd as va
    from oracle import oracle
    db = va.Virginian()
    name, cus = db.device()
    if cus == 0:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    db.gen_table("test", 1 << 20, 1 << 20, seed=a.seed, threads=16)
    q = db.prepare(CONFIG1_SQL)
    q.run()
    pair = [q.time(1, 1)[0] for _ in range(20)]
    check = verify_columns(q.fetch(), golden("config1") if a.seed == 1 else None)
    dropin = [_execute_ms(va, db, CONFIG1_SQL)[0] for _ in range(21)][1:]
    cpu = []
    for _ in range(20):
        t0 = time.perf_counter()
        oracle.select(db, CONFIG1_SQL)
        cpu.append((time.perf_counter() - t0) * 1e3)
    read_b = q.rows * q.read_bytes_per_row

    def stats(v):
        return {"best_ms": round(min(sorted(v)[:5]), 4), "median_ms": round(statistics.median(v), 4)}

    print(f"# config 1: {q.last_rows} of {q.rows} rows; resident launch median {statistics.median(pair):.3f} ms "
          f"({'specialised single pass' if q.kernel()['fused'] else 'interpreter launch pair'}), "
          f"virg_vm_execute median {statistics.median(dropin):.3f} ms, single-core CPU VM median "
          f"{statistics.median(cpu):.1f} ms", file=sys.stderr, flush=True)
    line = {
        "metric": "latency of the plain filter SELECT over one 1M-row tablet (resident query launch)",
        "value": round(statistics.median(pair), 4), "unit": "ms", "n_gpus": 1, "steps": 20, "warmup": 1,
        "ms_per_step": round(statistics.median(pair), 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "i32+f32", "data": "synthetic (seeded Philox generator, 7-column schema)",
        "verified": check["verified"],
        "config": {"workload": "plain filter SELECT (23 % selected) over one 1M-row tablet resident in HBM",
                   "rows": q.rows, "selected": q.last_rows, "read_bytes": read_b,
                   "kernel": q.kernel(), "launch": stats(pair), "dropin_virg_vm_execute": stats(dropin)},
        "roofline": {"bound": "hbm", "achieved": round(read_b / statistics.median(pair) / 1e6, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(read_b / statistics.median(pair) / 1e6 / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": {"value": round(statistics.median(cpu), 3), "unit": "ms", "cores": 1, "kind": "port",
                         "sample": "the same query on the same table, single-core VM, median of 20"},
    }
    print(json.dumps(line))
    q.close()
    db.close()
    return 0


# the reference's published mean query time for its 10-query suite
# (virginian_amd.workload.SUITE_QUERIES) over the 8M-row table of
# db/generate.c, on a GTX 570 / i7-960 (paper section 7, Table 1), in ms
SUITE_PUBLISHED_MS = {"cached_gpu": 14.0, "mapped_gpu": 29.0, "multi_core_cpu": 113.0, "single_core_cpu": 505.0}


def bench_suite(a) -> int:
    """The reference's own published benchmark on this GPU: the 10-query
    suite over an 8M-row table of its generator's distributions.  Per
    query: the specialised kernel with the table resident in HBM (HIP
    events; the reference's "cached GPU" time is its kernel time,
    src/test/timing.cc:74) and the interpreter's passes, the drop-in
    virg_vm_execute with the tablets in the device cache and with
    use_mmap (the host tablets streamed over the link: the reference's
    mapped mode), and the CPU VM (the oracle) on one core and on every
    core, each beside the paper's Table 1 figure.  value = the mean
    resident query time."""
    import statistics
    import virginian_amd as va
    from oracle import oracle
    from virginian_amd.workload import SUITE_QUERIES, suite_table
    rows = a.rows or 8_000_000
    db = va.Virginian()
    name, cus = db.device()
    if cus == 0:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    t0 = time.perf_counter()
    suite_table(db, rows, a.tablet_rows, a.seed)
    print(f"# the reference's suite table: {rows} rows in {a.tablet_rows}-row tablets, built in "
          f"{time.perf_counter() - t0:.1f}s on {name} ({cus} CUs)", file=sys.stderr, flush=True)
    per = []
    for i, sql in enumerate(SUITE_QUERIES):
        q = db.prepare(sql)
        q.run()
        ms, n = q.time(a.warmup, a.steps)
        fused = q.kernel()["fused"]
        read_b, write_b = q.rows * q.read_bytes_per_row, n * q.write_bytes_per_row
        q.close()
        qi = db.prepare(sql, flags=va.VG_QUERY_INTERP)
        qi.run()
        ms_i, _ = qi.time(a.warmup, a.steps)
        qi.close()
        _execute_ms(va, db, sql)
        warm = statistics.median(_execute_ms(va, db, sql)[0] for _ in range(3))
        db.use_mmap = True
        try:
            mapped = statistics.median(_execute_ms(va, db, sql)[0] for _ in range(3))
        finally:
            db.use_mmap = False
        per.append({"query": i, "selected": n, "selectivity": round(n / q.rows, 5), "kernel_ms": round(ms, 4),
                    "specialised": fused, "read_bytes": read_b, "write_bytes": write_b,
                    "interpreter_ms": round(ms_i, 4), "dropin_cached_ms": round(warm, 3),
                    "dropin_mapped_ms": round(mapped, 3)})
        print(f"# Q{i}: {n} rows ({n / q.rows:.2%}) {ms * 1e3:.1f} us resident ({'specialised' if fused else 'interpreter'})"
              f", interpreter {ms_i * 1e3:.1f} us; virg_vm_execute cached {warm:.2f} ms, mapped {mapped:.2f} ms",
              file=sys.stderr, flush=True)
    # the CPU VM on the same table: one pass of the suite on every core, one on one core
    threads = cpu_threads()
    t0 = time.perf_counter()
    for sql in SUITE_QUERIES:
        oracle.select(db, sql, threads=threads, ordered=False)
    cpu_multi = (time.perf_counter() - t0) * 1e3 / len(SUITE_QUERIES)
    t0 = time.perf_counter()
    for sql in SUITE_QUERIES:
        oracle.select(db, sql)
    cpu_single = (time.perf_counter() - t0) * 1e3 / len(SUITE_QUERIES)
    print(f"# CPU VM mean per query: {cpu_single:.1f} ms on one core, {cpu_multi:.1f} ms on {threads}",
          file=sys.stderr, flush=True)
    mean = statistics.mean(p["kernel_ms"] for p in per)
    alg = statistics.mean(p["read_bytes"] + p["write_bytes"] for p in per)
    means = {"resident_specialised": round(mean, 4),
             "resident_interpreter": round(statistics.mean(p["interpreter_ms"] for p in per), 4),
             "dropin_cached": round(statistics.mean(p["dropin_cached_ms"] for p in per), 3),
             "dropin_mapped": round(statistics.mean(p["dropin_mapped_ms"] for p in per), 3),
             "cpu_multi_core": round(cpu_multi, 2), "cpu_single_core": round(cpu_single, 2)}
    line = {
        "metric": "mean query time of the reference's 10-query filter SELECT suite (src/test/queries.h), 8M rows",
        "value": round(mean, 4), "unit": "ms", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(mean * len(SUITE_QUERIES), 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": round(mean / SUITE_PUBLISHED_MS["cached_gpu"], 6), "dtype": "i32+f32",
        "data": "synthetic (the reference generator's distributions, seeded)",
        "config": {"workload": "the reference's 10-query suite over its 8M-row test table resident in HBM",
                   "rows": rows, "tablet_rows": a.tablet_rows, "means_ms": means,
                   "published_ms": SUITE_PUBLISHED_MS,
                   "published_hardware": "GTX 570 (152 GB/s) and i7-960 (4 cores, 8 threads), paper section 7",
                   "cpu_threads": threads, "cpu_model": cpu_model(), "per_query": per},
        "roofline": {"bound": "hbm", "achieved": round(alg / mean / 1e6, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg / mean / 1e6 / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": {"value": round(cpu_multi, 2), "unit": "ms", "cores": threads, "kind": "port",
                         "sample": f"one pass of the suite on the same table, {threads} threads; one core: "
                                   f"{cpu_single:.1f} ms per query"},
    }
    print(json.dumps(line))
    db.close()
    return 0


CONFIG3_SQL = ("select id, uniformi, normalf20 from test "
               "where uniformf < 9.24 and normalf5 > 2.0 or normali20 > 10")


def bench_outofcore(a) -> int:
    """Out-of-core filter SELECT: a 2e9-row table in pinned host memory (1908
    tablets of 1M rows, the last partial); every step streams the
    referenced columns (28 B/row) through HBM.  value = link bytes / wall."""
    import virginian_amd as va
    rows = a.rows if a.rows else 2_000_000_000
    db = va.Virginian()
    name, cus = db.device()
    if cus == 0:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    h2d_peak, d2h_peak = db.link_probe(1 << 30, 3)
    print(f"# pinned copy: {h2d_peak:.1f} GB/s host-to-device, {d2h_peak:.1f} GB/s device-to-host",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    os.environ["VIRG_GEN_PROGRESS"] = str(1 << 28)
    # the system pins about 64 GiB at most: keep some of it for the staging
    # and result buffers; tablets beyond the limit are staged by host threads
    db.pin_limit(a.pin_limit_gib << 30)
    # the device table builder (the same bits as the host generator)
    db.gen_table("test", rows, a.tablet_rows, seed=a.seed, threads=16, device=not a.host_gen)
    pinned, pageable = db.host_memory()
    print(f"# generated {rows} rows ({'host' if a.host_gen else 'device'} builder) in "
          f"{time.perf_counter() - t0:.1f}s; tablets: {pinned / 2**30:.1f} GiB "
          f"pinned, {pageable / 2**30:.1f} GiB pageable", file=sys.stderr, flush=True)
    cap = rows // 8
    for _ in range(a.warmup):
        db.stream_select(CONFIG3_SQL, capacity=cap, batch_rows=a.batch_rows)
    link = secs = 0.0
    n = 0
    per = []
    last = None
    for _ in range(a.steps):
        last, st = db.stream_select(CONFIG3_SQL, capacity=cap, batch_rows=a.batch_rows)
        link += st.h2d_bytes + st.d2h_bytes
        secs += st.seconds
        n = st.rows_out
        per.append(round(st.seconds * 1e3, 1))
        print(f"# {st.rows_scanned} rows -> {st.rows_out} in {st.seconds:.3f}s, "
              f"{st.h2d_bytes / st.seconds / 1e9:.1f} GB/s columns in, {st.batches} batches",
              file=sys.stderr, flush=True)
    value = link / secs / 1e9
    # the last step's rows against the oracle's digest (outside the timed steps)
    check = verify_columns(last, golden("config3") if (rows, a.tablet_rows, a.seed) == (2_000_000_000, 1 << 20, 1)
                           else None) if last is not None else {"verified": None}
    del last
    print(f"# results against the oracle's digest: {check['verified']}", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu:
        from oracle import oracle
        threads = cpu_threads()
        sample_tabs = 200
        t1 = time.perf_counter()
        got, _ = oracle.select(db, CONFIG3_SQL, threads=threads, ordered=False,
                               max_tablets=sample_tabs)
        dt = time.perf_counter() - t1
        srows = min(rows, sample_tabs * a.tablet_rows)
        cpu = {"value": round(srows * 28 / dt / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
               "sample": f"first {sample_tabs} tablets ({srows} rows) of the same host table, {dt:.1f}s"}
    line = {
        "metric": "GB/s over the host link for out-of-core filter SELECT", "value": round(value, 2),
        "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(secs / a.steps * 1e3, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+i32", "data": "synthetic (seeded Philox generator)",
        "verified": check["verified"],
        "config": {"workload": "out-of-core filter SELECT over a 2e9-row 7-column table in pinned host "
                               "memory (1M-row tablets), referenced columns streamed through HBM",
                   "rows": rows, "tablet_rows": a.tablet_rows, "selected": n,
                   "link_bytes_per_step": round(link / a.steps), "step_ms": per,
                   "pinned_copy_gbs": {"h2d": round(h2d_peak, 1), "d2h": round(d2h_peak, 1)},
                   "tablet_bytes": {"pinned": pinned, "pageable": pageable}},
        "roofline": {"bound": "link", "achieved": round(value, 2), "peak": round(h2d_peak, 1),
                     "unit": "GB/s", "frac": round(value / h2d_peak, 4), "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    db.close()
    return 0


def _digest(q) -> str:
    import hashlib
    h = hashlib.sha1()
    for name, col in q.fetch().items():
        h.update(name.encode())
        h.update(col.tobytes())
    return h.hexdigest()


def _pct(xs: list[float], p: float) -> float:
    import numpy as np
    return float(np.percentile(np.asarray(xs), p))


def bench_batched(a) -> int:
    """Batched queries: 256 seeded random op-programs (4-40 ops, a
    uniformf < 100*s conjunct with s ~ U[0.01 %, 80 %]) over 16 resident
    tables of log-uniform 1e3-1e7 rows, run by the specialised batch (the
    programs compiled into one module at build time, one persistent launch
    taking (query, chunk) items from a device counter) and, for
    comparison, by the interpreter's tuned persistent batch and by plain
    per-query launches on one stream.  value = GB/s scanned by the
    specialised batch (the config-2 byte formula summed over the queries);
    queries/s and p50/p99 latency for all three."""
    import numpy as np
    import virginian_amd as va
    from virginian_amd import workload as W
    db = va.Virginian()
    name, cus = db.device()
    if cus == 0:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    rng = np.random.default_rng(a.seed)
    sizes = W.log_uniform_sizes(rng, va.VIRG_MAX_TABLES, 1e3, 1e7)
    tids = []
    for k, rows in enumerate(sizes):
        db.gen_table(f"t{k}", rows, a.tablet_rows, seed=a.seed + k, threads=16)
        tids.append(db.table_id(f"t{k}"))
    progs, qtab = [], []
    while len(progs) < a.queries:
        k = int(rng.integers(0, len(tids)))
        if a.programs and len(progs) >= a.programs:
            # the structure of an earlier program over another table
            base = progs[len(progs) % a.programs]
            progs.append([(o[0], tids[k], *o[2:]) if o[0] == va.OP_Table else o for o in base])
            qtab.append(k)
            continue
        p = W.random_program(rng, max_ops=40, selectivity=float(rng.uniform(0.0001, 0.8)), table=tids[k])
        if len(p) < 4:
            continue
        progs.append(p)
        qtab.append(k)
    qs = [db.prepare(p, flags=va.VG_QUERY_INTERP) for p in progs]
    scanned = sum(q.rows for q in qs)
    print(f"# {len(qs)} queries over {len(sizes)} tables ({sum(sizes)} rows), {scanned} rows scanned "
          f"per round; ops per program {min(map(len, progs))}-{max(map(len, progs))}",
          file=sys.stderr, flush=True)
    if a.pmc_run:
        # counter runs: every query's plain launches once, then one run of
        # the interpreter's batch (default configuration) and one of the
        # specialised batch, told apart by kernel name
        for q in qs:
            q.run()
        b = va.PersistentBatch(qs)
        rows_i, _ = b.run()
        b.close()
        print(f"# compiling the batch module ({len(qs)} programs)", file=sys.stderr, flush=True)
        jb = va.PersistentBatch(qs, jit=1)
        rows_j, _, flags = jb.results()
        jb.close()
        nbytes = sum(q.rows * q.read_bytes_per_row + r * q.write_bytes_per_row for q, r in zip(qs, rows_j))
        print(json.dumps({"pmc_run": "batched", "queries": len(qs), "rows_scanned": scanned,
                          "bytes_algorithmic": nbytes, "same_rows": rows_i == rows_j and not any(flags)}))
        return 0
    # plain: each query's launch pair, one after another
    plain_ms, plain_lat = va.plain_time(qs, a.warmup, a.steps)
    plain_rows = [q.sync() for q in qs]
    plain_dig = [_digest(q) for q in qs]
    print(f"# plain: {plain_ms:.3f} ms per round, p50 {_pct(plain_lat, 50):.3f} p99 {_pct(plain_lat, 99):.3f} ms",
          file=sys.stderr, flush=True)
    b = va.PersistentBatch(qs)
    t0 = time.perf_counter()
    cfg, tune_ms = b.tune(iters=3)
    cfg = b.config()
    print(f"# tuned in {time.perf_counter() - t0:.1f}s: {cfg} {tune_ms:.3f} ms", file=sys.stderr, flush=True)
    batch_ms = b.time(a.warmup, a.steps)
    lats = []
    for _ in range(a.steps):
        rows, lat = b.run()
        lats.append(lat)
    batch_lat = [float(np.median([l[i] for l in lats])) for i in range(len(qs))]
    same = rows == plain_rows and [_digest(q) for q in qs] == plain_dig
    print(f"# batch: {batch_ms:.3f} ms per run, p50 {_pct(batch_lat, 50):.3f} p99 {_pct(batch_lat, 99):.3f} ms, "
          f"results identical to plain: {same}", file=sys.stderr, flush=True)
    b.close()
    # specialised: the batch's programs compiled into one module, one
    # persistent launch taking (query, chunk) items in order
    t0 = time.perf_counter()
    print(f"# compiling the batch module ({len(qs)} programs)", file=sys.stderr, flush=True)
    jb = va.PersistentBatch(qs, jit=1)
    compile_s = time.perf_counter() - t0
    jit_ms = jb.time(a.warmup, a.steps)
    jlats = []
    for _ in range(a.steps):
        jrows, jlat, jflags = jb.results()
        jlats.append(jlat)
    jit_lat = [float(np.median([l[i] for l in jlats])) for i in range(len(qs))]
    jit_same = jrows == plain_rows and not any(jflags) and [_digest(q) for q in qs] == plain_dig
    print(f"# specialised batch: built in {compile_s:.1f}s, {jit_ms:.3f} ms per run, p50 {_pct(jit_lat, 50):.3f} "
          f"p99 {_pct(jit_lat, 99):.3f} ms, results identical to plain: {jit_same}", file=sys.stderr, flush=True)
    jb.close()
    same = same and jit_same
    interp_ms, interp_lat = batch_ms, batch_lat
    batch_ms, batch_lat = jit_ms, jit_lat
    nbytes = sum(q.rows * q.read_bytes_per_row + r * q.write_bytes_per_row for q, r in zip(qs, rows))
    value = nbytes / (batch_ms * 1e-3) / 1e9
    line = {
        "metric": "GB/s scanned by a batch of filter SELECT queries (persistent batch vs plain launches)",
        "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(batch_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+i32", "data": "synthetic (seeded Philox generator, seeded programs)",
        "config": {"workload": f"batched: {len(qs)} random op-programs (4-40 ops) over {len(sizes)} resident "
                               "tables of log-uniform 1e3-1e7 rows, selectivity U[0.01 %, 80 %]"
                               + (f", {a.programs} distinct structures" if a.programs else ""),
                   "queries": len(qs), "table_rows": sizes, "rows_scanned_per_round": scanned,
                   "bytes_per_round": nbytes, "pct_hbm_peak": round(value / HBM_PEAK_GBS * 100, 2),
                   "batched": {"kind": "one persistent launch of the batch's programs compiled together",
                               "build_s": round(compile_s, 1),
                               "ms": round(batch_ms, 4), "queries_per_s": round(len(qs) / batch_ms * 1e3, 1),
                               "p50_ms": round(_pct(batch_lat, 50), 4), "p99_ms": round(_pct(batch_lat, 99), 4)},
                   "interpreter_batch": {"tuned": cfg, "ms": round(interp_ms, 4),
                                         "queries_per_s": round(len(qs) / interp_ms * 1e3, 1),
                                         "p50_ms": round(_pct(interp_lat, 50), 4),
                                         "p99_ms": round(_pct(interp_lat, 99), 4)},
                   "plain": {"ms": round(plain_ms, 4), "queries_per_s": round(len(qs) / plain_ms * 1e3, 1),
                             "p50_ms": round(_pct(plain_lat, 50), 4), "p99_ms": round(_pct(plain_lat, 99), 4)},
                   "speedup": round(plain_ms / batch_ms, 3), "identical_results": same},
        "roofline": {"bound": "hbm", "achieved": round(value, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(value / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    for q in qs:
        q.close()
    db.close()
    return 0 if same else 3


def _program_bytes(ops: list, table_rows: int, rows_out: int) -> int:
    """Algorithmic bytes of one query: every referenced column read once in
    full, every selected row written once."""
    import virginian_amd as va
    width = [4, 8, 4, 8, 1]
    cols = {("key" if o[0] == va.OP_Rowid else o[2]) for o in ops if o[0] in (va.OP_Column, va.OP_Rowid)}
    rin = sum(8 if c == "key" else 4 for c in cols)
    rout = sum(width[o[1]] for o in ops if o[0] == va.OP_ResultColumn)
    return table_rows * rin + rows_out * rout


def bench_served(a) -> int:
    """Served queries with learned sizing: Poisson arrivals (seeded
    exponential gaps) of the batched mix over 10 resident tables of
    log-uniform 1e5-5e7 rows; continuous batching with admission by the
    HBM arena; results complete in HBM.  The predictor (3-layer MLP, 32
    hidden, f32) is trained on the GPU on per-tablet counts from two-pass
    runs of other programs, for the --quantile quantile of each tablet's
    selectivity (pinball loss), its margin calibrated on held-out programs
    for at most --rerun-target reruns; a second network trained on the
    mean absolute error gives the accuracy figure.  value = the
    predictor-sized server; reported beside the same stream sized by the
    independence estimate alone (its own margin, same target), the
    predictor-sized single pass, buffers for every table row and plain
    launches."""
    import numpy as np
    import virginian_amd as va
    from virginian_amd import predict as PR
    from virginian_amd import serve as S
    from virginian_amd import workload as W
    db = va.Virginian()
    name, cus = db.device()
    if cus == 0:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    rng = np.random.default_rng(a.seed)
    sizes = W.log_uniform_sizes(rng, 10, 1e5, 5e7)
    tids = []
    t0 = time.perf_counter()
    for k, rows in enumerate(sizes):
        db.gen_table(f"t{k}", rows, a.tablet_rows, seed=a.seed + k, threads=16)
        tids.append(db.table_id(f"t{k}"))
    stats = {t: PR.TableStats(db, t) for t in tids}
    print(f"# {len(sizes)} tables, {sum(sizes)} rows, generated in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)

    def programs(n):
        out = []
        while len(out) < n:
            p = W.random_program(rng, max_ops=40, selectivity=float(rng.uniform(0.0001, 0.8)),
                                 table=int(rng.choice(tids)))
            if len(p) >= 4:
                out.append(p)
        return out

    t0 = time.perf_counter()
    # a fifth of the training programs validate (early stopping)
    nval = max(1, a.train_programs // 5)
    x, t = PR.collect(db, programs(a.train_programs - nval), stats)
    xv, tv = PR.collect(db, programs(nval), stats)
    t1 = time.perf_counter()
    net = PR.MLP(seed=a.seed, loss="quantile", tau=a.quantile)
    fitq = net.fit(x, t, xv, tv)
    t2 = time.perf_counter()
    mnet = PR.MLP(seed=a.seed)
    fit = mnet.fit(x, t, xv, tv)
    loss = fit["validation_mae"]
    held = programs(a.calib_programs)
    hx, ht = PR.collect(db, held, stats)
    pred = net.forward(hx)
    preds, true_rows, rowsl, k = [], [], [], 0
    for ops in held:
        st = stats[PR.table_of(ops)]
        n = len(st.rows)
        preds.append(pred[k:k + n])
        true_rows.append(float((ht[k:k + n] * st.rows).sum()))
        rowsl.append(st.rows)
        k += n
    margin = PR.calibrate(preds, true_rows, rowsl, a.rerun_target)
    # the method without the network: the independence estimate (feature 3)
    # with its own margin, calibrated on the same held-out programs
    est, k = [], 0
    for ops in held:
        n = len(stats[PR.table_of(ops)].rows)
        est.append(hx[k:k + n, 3])
        k += n
    margin_est = PR.calibrate(est, true_rows, rowsl, a.rerun_target)
    mae = float(np.mean(np.abs(mnet.forward(hx) - ht)))
    mnet.close()
    mae_est = float(np.mean(np.abs(hx[:, 3] - ht)))
    mae_const = float(np.mean(np.abs(float(np.mean(t)) - ht)))
    print(f"# predictor: {len(x)} + {len(xv)} examples collected in {t1 - t0:.1f}s; sizing network "
          f"(quantile {a.quantile}) trained in {t2 - t1:.2f}s ({fitq['best_epoch']} of {fitq['epochs']} epochs); "
          f"point network validation MAE {loss:.4f} from {fit['validation_mae_start']:.4f}, held-out MAE "
          f"{mae:.4f} (independence estimate {mae_est:.4f}, mean selectivity {mae_const:.4f}); margin "
          f"{margin:.3f} (estimate alone {margin_est:.3f}) on {len(held)} held-out programs",
          file=sys.stderr, flush=True)

    n_req = int(a.rate * a.duration)
    arrivals = np.cumsum(rng.exponential(1.0 / a.rate, n_req))
    stream = [(float(ar), ops) for ar, ops in zip(arrivals, programs(n_req))]
    rows_of = {t: int(stats[t].rows.sum()) for t in tids}
    out = {}
    for mode in a.modes.split(","):
        # single-pass batches run one at a time: their tiles wait on each
        # other's published counts, and three such kernels side by side were
        # measured 20x slower than one
        srv = S.Server(db, tids, a.arena_gib << 30, mode=mode,
                       predictor=net if mode in ("sized", "single") else None,
                       margin=margin_est if mode == "estimate" else margin,
                       runners=a.single_runners if mode == "single" else a.runners, max_batch=a.max_batch,
                       keep_results=a.keep,
                       split=a.split, grid_div=a.grid_div, batch_line=a.batch_line, costly=int(a.costly),
                       host_arena=a.host_arena)
        reqs = [S.Request(arrival=ar, ops=ops, rid=i) for i, (ar, ops) in enumerate(stream)]
        try:
            summ = srv.run(reqs)
        finally:
            srv.close()
        nbytes = sum(_program_bytes(r.ops, rows_of[PR.table_of(r.ops)], r.rows) for r in reqs)
        summ["gbs"] = nbytes / summ["seconds"] / 1e9
        summ["rows"] = [r.rows for r in reqs]
        summ["bytes"] = nbytes
        out[mode] = summ
        print(f"# {mode}: {summ['requests']} requests in {summ['seconds']:.3f}s, "
              f"{summ['queries_per_s']:.0f} q/s, p50 {summ['p50_ms']:.2f} ms, p90 {summ['p90_ms']:.2f} ms, "
              f"p99 {summ['p99_ms']:.2f} ms, "
              f"reruns {summ['reruns']}, peak arena {summ['peak_arena_bytes'] / 2**20:.0f} MiB, "
              f"{summ['gbs']:.1f} GB/s; host {summ['host_seconds']}", file=sys.stderr, flush=True)
    net.close()
    rows_seen = [o["rows"] for o in out.values()]
    same = all(r == rows_seen[0] for r in rows_seen)

    def brief(m):
        return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in out[m].items() if k != "rows"}

    def ratio(num, den, key, inv=False):
        if num not in out or den not in out:
            return None
        x, y = out[num][key], out[den][key]
        return round((y / x) if inv else (x / max(y, 1e-12)), 3)

    head = "sized" if "sized" in out else next(iter(out))
    s1 = out[head]
    names = {"sized": "sized_two_pass", "estimate": "sized_by_estimate", "single": "sized_single_pass",
             "two-pass": "two_pass_every_row", "plain": "plain_unbatched"}
    cfg = {"workload": f"served: Poisson arrivals at {a.rate:.0f} q/s for {a.duration:.1f} s of random "
                       "op-programs (4-40 ops) over 10 resident tables of log-uniform 1e5-5e7 rows",
           "table_rows": sizes, "requests": n_req, "arena_gib": a.arena_gib,
           "server": {"runners": a.runners, "max_batch": a.max_batch, "split": a.split, "grid_div": a.grid_div,
                      "batch_line": a.batch_line, "costly": a.costly, "keep": a.keep, "host_arena": a.host_arena},
           "predictor": {"examples": int(len(x)), "train_s": round(t2 - t1, 3), "quantile": a.quantile,
                         "point_network_validation_mae": round(loss, 6),
                         "heldout_mae": round(mae, 4), "independence_estimate_mae": round(mae_est, 4),
                         "mean_selectivity_mae": round(mae_const, 4), "margin": margin,
                         "estimate_margin": margin_est}}
    for m in out:
        cfg[names.get(m, m)] = brief(m)
    cfg["rerun_rate"] = round(s1["reruns"] / n_req, 4)
    for m in ("two-pass", "sized", "estimate", "single"):
        cfg[f"{m.replace('-', '_')}_vs_plain"] = {
            "throughput": ratio(m, "plain", "queries_per_s"), "p50": ratio(m, "plain", "p50_ms", inv=True),
            "p90": ratio(m, "plain", "p90_ms", inv=True), "p99": ratio(m, "plain", "p99_ms", inv=True)}
    cfg["gain_vs_two_pass"] = {"throughput": ratio("sized", "two-pass", "queries_per_s"),
                               "p99": ratio("sized", "two-pass", "p99_ms", inv=True),
                               "reserved_bytes": ratio("sized", "two-pass", "reserved_bytes", inv=True)}
    cfg["gain_vs_estimate"] = {"reserved_bytes": ratio("sized", "estimate", "reserved_bytes", inv=True),
                               "reruns": [out[m]["reruns"] for m in ("sized", "estimate") if m in out]}
    cfg["identical_row_counts"] = same
    line = {
        "metric": "GB/s scanned by served filter SELECT queries (learned sizing vs buffers for every row)",
        "value": round(s1["gbs"], 2), "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(s1["seconds"] * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+i32", "data": "synthetic (seeded Philox generator, seeded programs)",
        "config": cfg,
        "roofline": {"bound": "hbm", "achieved": round(s1["gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(s1["gbs"] / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    db.close()
    return 0 if same else 3


def _execute_ms(va, db, sql: str) -> tuple[float, int]:
    """One virg_vm_execute (-> virg_vm_gpu): host tablets in, results in host
    result tablets; wall-clock ms and result rows."""
    vm = db.compile(sql)
    try:
        t0 = time.perf_counter()
        if va.lib.virg_vm_execute(db.ptr, vm) != va.VIRG_SUCCESS:
            raise RuntimeError("virg_vm_execute failed")
        ms = (time.perf_counter() - t0) * 1e3
        n = C.c_ulonglong(0)
        va.lib.virg_result_columns(db.ptr, vm, None, C.byref(n))
        return ms, n.value
    finally:
        db.free_vm(vm)


def dropin_times(db, per: list, a) -> dict:
    """The drop-in boundary, virg_vm_execute -> virg_vm_gpu, which hands over
    host buffers: warm (tablets already in the device tablet cache; the
    results cross the host link into result tablets), median of 3 per query;
    and cold for the 10 % query on a fresh instance (every tablet crosses
    the link first).  Reported beside value, never as value."""
    import statistics
    import virginian_amd as va
    warm = []
    for p in per:
        sql = sql_for(p["threshold"])
        _execute_ms(va, db, sql)
        runs = [_execute_ms(va, db, sql) for _ in range(3)]
        ms = statistics.median(r[0] for r in runs)
        warm.append({"threshold": p["threshold"], "ms": round(ms, 3), "rows": runs[0][1],
                     "gbs_scanned": round(p["read_bytes"] / ms / 1e6, 1),
                     "result_bytes": p["write_bytes"]})
        print(f"# drop-in virg_vm_execute T={p['threshold']}: {ms:.2f} ms with results into host "
              f"result tablets ({p['read_bytes'] / ms / 1e6:.0f} GB/s scanned)", file=sys.stderr, flush=True)
    fresh = va.Virginian()
    try:
        fresh.gen_table("test", a.rows, a.tablet_rows, seed=a.seed, threads=16)
        ms, n = _execute_ms(va, fresh, sql_for(10.0))
    finally:
        fresh.close()
    rb = per[1]["read_bytes"]
    print(f"# drop-in cold (tablets uploaded first) T=10.0: {ms:.1f} ms", file=sys.stderr, flush=True)
    return {"warm": warm, "cold_t10": {"ms": round(ms, 2), "rows": n, "gbs_scanned": round(rb / ms / 1e6, 1)}}


PMC_RUNS = 3


def pmc_run(va, db, queries) -> int:
    """Counter runs (rocprofv3 --pmc): each query's kernel PMC_RUNS times
    in threshold order, one at a time, so the dispatches can be told apart
    by their order; then the interpreter's count and write passes on the
    same queries, PMC_RUNS times each (the kernels before specialisation)."""
    for q in queries:
        for _ in range(PMC_RUNS):
            q.run()
    interp = []
    for t in THRESHOLDS:
        qi = db.prepare(sql_for(t), flags=va.VG_QUERY_INTERP)
        for _ in range(PMC_RUNS):
            qi.run()
        interp.append(qi.last_rows)
        qi.close()
    print(json.dumps({"pmc_run": [q.last_rows for q in queries], "interpreter": interp,
                      "runs_per_query": PMC_RUNS}))
    return 0


def summarize_counters(outdir: str, dest: str, workload: str = "batched") -> int:
    """Per-kernel means of every counter in the rocprofv3 --pmc runs under
    outdir (one directory per pass), for a workload's counter runs
    (--workload W --pmc-run: the batched workload, or the resident one):
    dispatches, and per dispatch and per run the counter values
    (FETCH_SIZE and WRITE_SIZE in KiB)."""
    wl = "--workload batched " if workload == "batched" else ""
    tag = "batched" if workload == "batched" else "config2"
    import csv
    from collections import defaultdict
    from pathlib import Path as P
    acc = defaultdict(lambda: defaultdict(list))
    for f in sorted(P(outdir).glob("**/*counter_collection.csv")):
        for r in csv.DictReader(open(f)):
            k = r["Kernel_Name"].split("(")[0]
            acc[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
    kernels = {}
    for k, cs in sorted(acc.items()):
        kernels[k] = {c: {"dispatches": len(v), "sum": round(sum(v)), "mean": round(sum(v) / len(v), 1)}
                      for c, v in sorted(cs.items())}
    doc = {"method": "rocprofv3 --pmc, one pass per counter group; per kernel: dispatches, sum over the run "
                     "and mean per dispatch (FETCH_SIZE / WRITE_SIZE in KiB, uncalibrated)",
           "commands": [
               "cd /tmp && export TMPDIR=/tmp",
               "rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_SMEM SQ_INSTS_LDS SQ_WAVES SQ_WAVE_CYCLES "
               f"SQ_BUSY_CYCLES SQ_WAIT_ANY --output-format csv -d <out>/sq -o {tag} -- python3 -u <repo>/bench.py "
               f"{wl}--pmc-run",
               f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d <out>/fetch -o {tag} -- python3 -u "
               f"<repo>/bench.py {wl}--pmc-run",
               f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d <out>/write -o {tag} -- python3 -u "
               f"<repo>/bench.py {wl}--pmc-run",
               f"python3 <repo>/bench.py {wl}--summarize-counters <out> {dest}"],
           "kernels": kernels,
           "hardware": "MI355X (gfx950, 256 CUs)"}
    P(dest).write_text(json.dumps(doc, indent=1) + "\n")
    print(json.dumps(doc, indent=1))
    return 0


def summarize_pmc(outdir: str, dest: str) -> int:
    """Per-launch HBM bytes from two counter runs of --pmc-run (FETCH_SIZE
    and WRITE_SIZE, in KiB per dispatch) and of --copy-probe 256 (the same
    counters on device copies of known size, which calibrate them), written
    as the traffic summary bench.py reads."""
    import csv
    from pathlib import Path as P

    def rows(counter, name):
        f = next(P(outdir).glob(f"pmc_{counter}/**/{name}_counter_collection.csv"))
        out = []
        for r in csv.DictReader(open(f)):
            if r["Counter_Name"] == counter:
                out.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
        return sorted(out)

    cal = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        copies = [v for _, k, v in rows(c, "copy") if "copy" in k.lower()]
        cal[c] = sorted(copies)[len(copies) // 2] / (256 * 1024)
    per = {c: [v * 1024 / cal[c] for _, k, v in rows(c, "config2") if k.startswith("vgx_main")]
           for c in ("FETCH_SIZE", "WRITE_SIZE")}
    n = len(THRESHOLDS)
    assert all(len(v) == n * (1 + PMC_RUNS) for v in per.values()), {k: len(v) for k, v in per.items()}
    queries = []
    for i, t in enumerate(THRESHOLDS):
        # the first dispatch of each query is bench.py's q.run() before the counter runs
        idx = [n + i * PMC_RUNS + j for j in range(PMC_RUNS)]
        rd = sum(per["FETCH_SIZE"][k] for k in idx) / PMC_RUNS
        wr = sum(per["WRITE_SIZE"][k] for k in idx) / PMC_RUNS
        queries.append({"threshold": t, "read_bytes": round(rd), "write_bytes": round(wr),
                        "bytes_per_launch": round(rd + wr)})
    doc = {"calibration": {k: round(v, 5) for k, v in cal.items()},
           "method": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per dispatch) / counter KiB per true KiB "
                     "measured on a 256 MiB device copy",
           "kernel": "vgx_main (the query's specialised single-pass kernel)",
           "queries": queries,
           "bytes_per_launch_mean": round(sum(q["bytes_per_launch"] for q in queries) / n),
           "commands": [
               "cd /tmp && export TMPDIR=/tmp",
               "rocprofv3 --pmc FETCH_SIZE --output-format csv -d <out>/pmc_FETCH_SIZE -o config2 -- "
               "python3 -u <repo>/bench.py --pmc-run",
               "rocprofv3 --pmc FETCH_SIZE --output-format csv -d <out>/pmc_FETCH_SIZE -o copy -- "
               "python3 -u <repo>/bench.py --copy-probe 256",
               "rocprofv3 --pmc WRITE_SIZE --output-format csv -d <out>/pmc_WRITE_SIZE -o config2 -- "
               "python3 -u <repo>/bench.py --pmc-run",
               "rocprofv3 --pmc WRITE_SIZE --output-format csv -d <out>/pmc_WRITE_SIZE -o copy -- "
               "python3 -u <repo>/bench.py --copy-probe 256",
               "python3 <repo>/bench.py --summarize-pmc <out> profiles/traffic_config2_64M.json"],
           "attribution": f"dispatches of vgx_main in order: one run per query, then {PMC_RUNS} per query "
                          "in threshold order; counter KiB x 1024 / calibration = bytes",
           "hardware": "MI355X (gfx950, 256 CUs)"}
    P(dest).write_text(json.dumps(doc, indent=1) + "\n")
    print(json.dumps(doc, indent=1))
    return 0


GOLDEN = ROOT / "tests" / "golden" / "workload_digests.json"


def golden(key: str) -> dict | None:
    """The oracle's digests of a workload (tests/golden/make_workload_digests.py)."""
    if not GOLDEN.exists():
        return None
    return json.loads(GOLDEN.read_text()).get(key)


def verify_columns(cols: dict, want: dict | None) -> dict:
    """Compare a result (column arrays in order) with its recorded digest."""
    from virginian_amd.verify import digest
    got = {"selected": len(next(iter(cols.values()))) if cols else 0, "digest": digest(list(cols.values()))}
    if want is None:
        return {**got, "verified": None}
    return {**got, "verified": got["selected"] == want["selected"] and got["digest"] == want["digest"]}


def measured_traffic(workload: str) -> dict | None:
    """HBM traffic per launch pair recorded from counter runs, if present."""
    p = ROOT / "profiles" / f"traffic_{workload}.json"
    if not p.exists():
        return None
    return json.loads(p.read_text())


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=0, help="table rows (default: the workload's)")
    ap.add_argument("--tablet-rows", type=int, default=1 << 20)
    ap.add_argument("--host-gen", action="store_true",
                    help="outofcore: build the table with the host generator instead of the device builder")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pin-limit-gib", type=int, default=52,
                    help="out-of-core: pin at most this much tablet memory")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB")
    ap.add_argument("--workload", choices=["resident", "outofcore", "batched", "served", "config1", "suite"],
                    default="resident")
    ap.add_argument("--rate", type=float, default=2000.0, help="served: arrivals per second")
    ap.add_argument("--duration", type=float, default=2.0, help="served: seconds of arrivals")
    ap.add_argument("--arena-gib", type=int, default=8, help="served: HBM for result buffers")
    ap.add_argument("--rerun-target", type=float, default=0.005,
                    help="served: rerun rate the predictor's margin is calibrated for")
    ap.add_argument("--train-programs", type=int, default=1600, help="served: predictor training runs")
    ap.add_argument("--calib-programs", type=int, default=600, help="served: held-out runs for the margins")
    ap.add_argument("--quantile", type=float, default=0.9, help="served: the sizing network's quantile")
    ap.add_argument("--runners", type=int, default=3, help="served: runs in flight at once (streams)")
    ap.add_argument("--max-batch", type=int, default=64, help="served: requests per run")
    ap.add_argument("--single-runners", type=int, default=1, help="served: runners of the single-pass server")
    ap.add_argument("--split", type=int, default=0, help="served: 1 = one launch group per batch")
    ap.add_argument("--grid-div", type=int, default=0, help="served: batches launch 1/grid-div of the grid")
    ap.add_argument("--batch-line", type=int, default=16, help="served: waiting requests from which runs batch")
    ap.add_argument("--keep", action="store_true", help="served: each result to host memory before it counts done")
    ap.add_argument("--host-arena", action="store_true", help="served: kernels write results into pinned host memory")
    ap.add_argument("--costly", type=float, default=1e8, help="served: rows x ops of a costly request")
    ap.add_argument("--modes", default="sized,estimate,single,two-pass,plain", help="served: servers to run")
    ap.add_argument("--queries", type=int, default=256, help="batched: queries per batch")
    ap.add_argument("--programs", type=int, default=0,
                    help="batched: distinct program structures (0: every query its own)")
    ap.add_argument("--batch-rows", type=int, default=8 << 20)
    ap.add_argument("--pmc-run", action="store_true",
                    help="counter runs: each resident query's kernel a few times, one at a time; "
                         "batched: plain launches, the interpreter's batch and the specialised batch once")
    ap.add_argument("--summarize-counters", nargs=2, metavar=("OUTDIR", "DEST"),
                    help="per-kernel means of the counters in the batched workload's counter runs")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("OUTDIR", "DEST"),
                    help="traffic summary from --pmc-run counter runs (no GPU needed)")
    a = ap.parse_args()
    if a.summarize_pmc:
        return summarize_pmc(*a.summarize_pmc)
    if a.summarize_counters:
        return summarize_counters(*a.summarize_counters, workload=a.workload)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the executor on one GPU; figures for several GPUs are "
              "not produced (tablets shard across GPUs by design, but no multi-GPU path is "
              "launched)", file=sys.stderr)
        return 2
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return 0
    if a.workload == "outofcore":
        return bench_outofcore(a)
    if a.workload == "config1":
        return bench_config1(a)
    if a.workload == "suite":
        return bench_suite(a)
    if a.workload == "batched":
        return bench_batched(a)
    if a.workload == "served":
        return bench_served(a)
    if not a.rows:
        a.rows = 64 << 20

    import virginian_amd as va

    db = va.Virginian()
    name, cus = db.device()
    if cus == 0:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    t0 = time.perf_counter()
    db.gen_table("test", a.rows, a.tablet_rows, seed=a.seed, threads=16)
    print(f"# generated {a.rows} rows in {time.perf_counter() - t0:.1f}s on {name} ({cus} CUs)",
          file=sys.stderr, flush=True)

    queries = [db.prepare(sql_for(t)) for t in THRESHOLDS]
    for q in queries:
        q.run()
    kinfo = queries[0].kernel()
    if a.pmc_run:
        return pmc_run(va, db, queries)

    # per query: kernel time back to back (HIP events on the query's
    # stream), beside the interpreter's count + write passes for the same
    # query
    per = []
    for t, q in zip(THRESHOLDS, queries):
        ms, n = q.time(a.warmup, a.steps)
        qi = db.prepare(sql_for(t), flags=va.VG_QUERY_INTERP)
        qi.run()
        ms_i, _ = qi.time(a.warmup, a.steps)
        ms_c, ms_w, _ = qi.time_passes(a.warmup, a.steps)
        qi.close()
        read_b = q.rows * q.read_bytes_per_row
        write_b = n * q.write_bytes_per_row
        per.append({"threshold": t, "selected": n, "selectivity": n / q.rows, "kernel_ms": ms,
                    "read_bytes": read_b, "write_bytes": write_b,
                    "gbs_scanned": read_b / ms / 1e6, "gbs_total": (read_b + write_b) / ms / 1e6,
                    "interpreter_ms": ms_i, "interpreter_count_pass_ms": ms_c,
                    "interpreter_write_pass_ms": ms_w})
        print(f"# T={t}: {n} rows ({n / q.rows:.4%}) {ms:.3f} ms/launch "
              f"({'specialised single pass' if q.kernel()['fused'] else 'interpreter'}), "
              f"{read_b / ms / 1e6:.0f} GB/s scanned, {(read_b + write_b) / ms / 1e6:.0f} GB/s algorithmic; "
              f"interpreter {ms_i:.3f} ms (count {ms_c:.3f} + write {ms_w:.3f})", file=sys.stderr, flush=True)

    dropin = dropin_times(db, per, a)

    # whole job: K steps of the 4-query sweep, one stream per query
    def sync_all():
        for q in queries:
            n = C.c_ulonglong()
            if va.lib.vg_query_sync(q.h, C.byref(n)) != va.VIRG_SUCCESS:
                raise RuntimeError("query failed")

    for _ in range(a.warmup):
        for q in queries:
            va.lib.vg_query_launch(q.h)
        sync_all()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        for q in queries:
            if va.lib.vg_query_launch(q.h) != va.VIRG_SUCCESS:
                raise RuntimeError("launch failed")
    sync_all()
    wall = time.perf_counter() - t0
    scanned = sum(p["read_bytes"] for p in per) * a.steps
    value = scanned / wall / 1e9

    # the rows the timed launches produced, against the oracle's digests
    # (outside the timed region)
    g = golden("config2") if (a.rows, a.tablet_rows, a.seed) == (64 << 20, 1 << 20, 1) else None
    checks = []
    for t, q in zip(THRESHOLDS, queries):
        q.sync()
        want = next((x for x in g["queries"] if x["threshold"] == t), None) if g else None
        checks.append({"threshold": t, **verify_columns(q.fetch(), want)})
    verified = all(c["verified"] for c in checks) if g else None
    print(f"# results against the oracle's digests: {'verified' if verified else verified}",
          file=sys.stderr, flush=True)

    mean_bytes = sum(p["read_bytes"] + p["write_bytes"] for p in per) / len(per)
    mean_ms = sum(p["kernel_ms"] for p in per) / len(per)
    achieved = mean_bytes / mean_ms / 1e6
    tr = measured_traffic("config2_64M")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": tr.get("bytes_per_launch_mean") if tr else None,
                "kernel": ("vgx_main: the query's specialised single-pass kernel" if kinfo["fused"]
                           else "vg_count_kernel + vg_write_kernel (one query)"),
                "algorithmic_bytes_per_launch": round(mean_bytes)}

    cpu = None
    if not a.no_cpu:
        # the T=10 query over the same rows, in the GPU's 1M-row tablets and
        # in the reference's 8 MB tablets
        q10 = queries[1]
        rb = q10.rows * q10.read_bytes_per_row
        r = cpu_baseline(db, sql_for(10.0), a.cpu_seconds)
        db.gen_table("test8mb", a.rows, 0, seed=a.seed, threads=cpu_threads())
        r8 = cpu_baseline(db, sql_for(10.0), a.cpu_seconds, table="test8mb")
        cal = cpu_calibration()
        ratio = cal["queries"]["config2_T10"]["multi"]["restatement_over_reference"] if cal else None
        value_cpu = rb * r["reps"] / r["seconds"] / 1e9
        cpu = {"value": round(value_cpu, 3), "unit": "GB/s", "cores": r["threads"], "kind": "port",
               "sample": f"the T=10 query over the same {q10.rows}-row table in 1M-row tablets, results "
                         f"materialised, {r['threads']} threads ({len(os.sched_getaffinity(0))} cores visible, "
                         f"{cpu_model()}), {r['reps']} full passes in {r['seconds']:.1f}s "
                         f"({q10.rows * r['reps'] / r['seconds']:.3e} rows/s)",
               "tablets_8mb": {"value": round(rb * r8["reps"] / r8["seconds"] / 1e9, 3), "unit": "GB/s",
                               "sample": f"the same query over the same rows in the reference's 8 MB tablets, "
                                         f"{r8['reps']} passes in {r8['seconds']:.1f}s"},
               "calibration": None if cal is None else {
                   "restatement_over_reference_time": ratio,
                   "reference_vm_estimate_gbs": round(value_cpu * ratio, 3),
                   "measured": f"{cal['rows']} rows, {cal['threads']} threads, {cal['cpu']} "
                               "(tests/golden/cpu_calibration.json)"}}

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "GB/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+i32",
        "data": "synthetic (seeded Philox generator, 7-column benchmark schema)",
        "verified": verified,
        "config": {"workload": "filter+math SELECT, 64M-row 7-column table resident in HBM in 1M-row "
                               "tablets, selectivity sweep 0.1/10/50/90 %",
                   "rows": a.rows, "tablet_rows": a.tablet_rows, "queries_per_step": len(queries),
                   "pct_hbm_peak": round(value / HBM_PEAK_GBS * 100, 2),
                   "kernel": {k: (round(v, 1) if isinstance(v, float) else v) for k, v in kinfo.items()},
                   "per_query": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in p.items()}
                                 for p in per],
                   "dropin_virg_vm_execute": dropin,
                   "verification": {"fixture": "tests/golden/workload_digests.json (config2)",
                                    "queries": checks}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    for q in queries:
        q.close()
    db.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
