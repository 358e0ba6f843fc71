"""Serving loop: queries arrive over time and are run with continuous
batching against resident tables.

Each step of the loop (one host thread):
  1. arrived requests join the waiting line;
  2. while a batch runner is free, the waiting queries are admitted in
     arrival order as long as their output buffers fit in the HBM arena
     (admission by footprint); each is prepared against its table's shared
     tablet descriptors and given buffers sized by the result-size
     predictor (modes "single" and "sized") or for every row of its table
     (mode "two-pass", the method without a predictor); the admitted queries go to the GPU
     as one batch on the runner's stream, while earlier batches are still
     running on theirs, so their work overlaps on the device;
  3. every finished batch is collected: results are copied to the host
     (the request is then complete), a query whose result did not fit is
     run again in two passes with the exact size the single pass counted,
     and its buffers return to the arena.
Latency is measured from a request's arrival to its result on the host.
No counterpart in the reference (its queries run one at a time,
src/vm/execute.c:28-102).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

import ctypes as C

from . import (NP_TYPE, OP_ResultColumn, VG_QUERY_LIGHT, VG_QUERY_NO_OUTPUT, VIRG_SUCCESS, PersistentBatch,
               VirginianError, lib, make_ops)
from . import predict as PR


class Arena:
    """First-fit allocator over one block of device memory."""

    def __init__(self, db, nbytes: int):
        self.db = db
        self.base = db.device_alloc(nbytes)
        self.size = nbytes
        self.free = [(0, nbytes)]          # (offset, size), sorted
        self.used = 0

    def alloc(self, nbytes: int) -> int | None:
        n = (nbytes + 255) & ~255
        for i, (off, sz) in enumerate(self.free):
            if sz >= n:
                if sz == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (off + n, sz - n)
                self.used += n
                return off
        return None

    def release(self, off: int, nbytes: int) -> None:
        n = (nbytes + 255) & ~255
        self.used -= n
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, s in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + s)
            else:
                merged.append((o, s))
        self.free = merged

    def largest(self) -> int:
        return max((sz for _, sz in self.free), default=0)

    def close(self) -> None:
        if self.base:
            self.db.device_free(self.base)
            self.base = 0


@dataclass
class Request:
    arrival: float                  # seconds after the start
    ops: list
    rid: int = 0
    ready: float = -1.0             # seconds after the start
    rows: int = 0
    rerun: bool = False
    cap: int = -1                   # output rows reserved (set at admission)
    result: dict = field(default_factory=dict)
    wire: tuple = ()                # the op-program as submitted (virg_op array)
    widths: tuple = ()              # bytes per value of each result column
    need: int = -1                  # arena bytes for cap rows (set with cap)

    def __post_init__(self):
        if not self.widths:
            self.widths = tuple(_WIDTH[o[1]] for o in self.ops if o[0] == OP_ResultColumn)
        if not self.wire:
            arr, n, keep = make_ops(self.ops)
            names = [o[4] if isinstance(o[4], str) else f"c{k}"
                     for k, o in enumerate([o for o in self.ops if o[0] == OP_ResultColumn])]
            types = [o[1] for o in self.ops if o[0] == OP_ResultColumn]
            self.wire = (arr, n, keep, names, types)


class _Query:
    """A served query: the handle of a light vg_query and what fetching
    its result needs."""

    def __init__(self, db, r: Request, cap: int, outs: list[int]):
        arr, n, _, self.names, self.types = r.wire
        err = C.create_string_buffer(256)
        self.h = lib.vg_query_create(db.ptr, arr, n, cap, VG_QUERY_LIGHT | VG_QUERY_NO_OUTPUT, err, 256)
        if not self.h:
            raise VirginianError(err.value.decode() or "vg_query_create failed")
        ptrs = (C.c_void_p * max(len(outs), 1))(*outs)
        lib.vg_query_set_outputs(self.h, ptrs, cap)
        self.last_rows = 0

    def fetch(self) -> dict:
        outs = [np.empty(self.last_rows, dtype=NP_TYPE[t]) for t in self.types]
        ptrs = (C.c_void_p * max(len(outs), 1))(*[o.ctypes.data for o in outs])
        if lib.vg_query_fetch(self.h, ptrs, self.last_rows) != VIRG_SUCCESS:
            raise VirginianError("vg_query_fetch failed")
        return dict(zip(self.names, outs))

    def close(self) -> None:
        if self.h:
            lib.vg_query_free(self.h)
            self.h = None


# vg_ctrl overflow bits (csrc/gpu/vg_prog.h): a single-pass wait gave up
VG_OVF_WAIT = 2

# bytes per value of each virg_t (INT, INT64, FLOAT, DOUBLE, CHAR)
_WIDTH = [4, 8, 4, 8, 1]


def _widths(ops: list) -> list[int]:
    return [_WIDTH[o[1]] for o in ops if o[0] == OP_ResultColumn]


class Server:
    """mode "single": predictor-sized buffers, single pass; "sized":
    predictor-sized buffers, count and write launches (the write pass flags
    what does not fit); "two-pass": buffers for every row of the table,
    count and write launches (the method without a predictor); "plain":
    no batching, each request run alone as it arrives."""

    def __init__(self, db, table_ids: list[int], arena_bytes: int, mode: str = "single", predictor=None,
                 margin: float = 0.05, runners: int = 3, max_batch: int = 64, tiles_per_item: int = 8,
                 keep_results: bool = True):
        self.db = db
        self.mode = mode
        self.stats = {t: PR.TableStats(db, t) for t in table_ids}
        self.net = predictor
        self.margin = margin
        self.arena = Arena(db, arena_bytes)
        single = 1 if mode == "single" else 0
        self.runners = [PersistentBatch([], tiles_per_item=tiles_per_item, single=single, direct=1, db=db)
                        for _ in range(runners)]
        # results that did not fit run again here, in two passes at their exact size
        self.rerun_runner = PersistentBatch([], tiles_per_item=tiles_per_item, single=0, direct=1, db=db)
        self.max_batch = max_batch
        self.keep = keep_results
        self.peak_bytes = 0
        self.reserved = 0                   # result bytes reserved, over all admissions
        # host seconds spent per step of the loop, and batches launched
        self.prof = {"admit": 0.0, "launch": 0.0, "collect": 0.0, "idle": 0.0, "batches": 0}

    def close(self) -> None:
        for r in self.runners + [self.rerun_runner]:
            r.close()
        self.arena.close()

    # -- admission -------------------------------------------------------
    def _capacities(self, reqs: list[Request]) -> list[int]:
        if self.mode == "two-pass":
            return [int(self.stats[PR.table_of(r.ops)].rows.sum()) for r in reqs]
        feats = [PR.program_features_ops(r.wire[0], r.wire[1], self.stats[PR.table_of(r.ops)]) for r in reqs]
        # on the host: the batches in flight hold every compute unit, and a
        # device forward pass would wait for them
        pred = self.net.forward_host(np.concatenate(feats)) if self.net is not None else None
        caps, k = [], 0
        for r, f in zip(reqs, feats):
            st = self.stats[PR.table_of(r.ops)]
            p = pred[k:k + len(f)] if pred is not None else f[:, 3]
            k += len(f)
            caps.append(PR.capacity(p, st.rows, self.margin))
        return caps

    def _admit(self, waiting: deque) -> tuple[list, tuple | None]:
        """Queries for one batch, in arrival order while their buffers fit
        in one free block of the arena: prepared, with their buffers."""
        cand = [waiting[i] for i in range(min(len(waiting), self.max_batch))]
        todo = [r for r in cand if r.cap < 0]
        for r, c in zip(todo, self._capacities(todo) if todo else []):
            r.cap = c
            r.need = -1
        room = self.arena.largest()
        take, total = [], 0
        for r in cand:
            if r.need < 0:
                r.need = sum((max(r.cap, 1) * w + 255) & ~255 for w in r.widths)
            need = r.need
            if total + need > room:
                break                       # admission is in arrival order
            take.append((r, need))
            total += need
            self.reserved += need
        if not take:
            return [], None
        base = self.arena.alloc(total)
        batch, off = [], base
        for r, need in take:
            outs = []
            for w in r.widths:
                outs.append(self.arena.base + off)
                off += (max(r.cap, 1) * w + 255) & ~255
            batch.append((r, _Query(self.db, r, max(r.cap, 1), outs)))
            waiting.popleft()
        self.peak_bytes = max(self.peak_bytes, self.arena.used)
        return batch, (base, total)

    # -- completion ------------------------------------------------------
    def _finish(self, runner, batch, block, t0: float, reruns: deque) -> int:
        """Collect a finished batch; returns the requests completed."""
        rows, _, flags = runner.collect()
        done = 0
        for (r, q), n, f in zip(batch, rows, flags):
            if f:
                # did not fit: again, in two passes, at the exact size
                # counted -- unless a single-pass wait gave up, which leaves
                # no usable count: then at the table's size
                r.cap = int(n) if not (f & VG_OVF_WAIT) else int(self.stats[PR.table_of(r.ops)].rows.sum())
                r.need = -1
                r.rerun = True
                reruns.append(r)
            else:
                q.last_rows = int(n)
                if self.keep:
                    r.result = q.fetch()
                r.ready = time.perf_counter() - t0
                r.rows = int(n)
                done += 1
        # the runner lets go of the queries before they are freed
        runner.reload([])
        for _, q in batch:
            q.close()
        self.arena.release(*block)
        return done

    def _run_plain(self, reqs: list[Request]) -> None:
        """The plain way: each request, in arrival order, prepared and run
        on its own (count and write launches, buffers for every row of its
        table), waited for, then the next."""
        t0 = time.perf_counter()
        for r in sorted(reqs, key=lambda r: r.arrival):
            while time.perf_counter() - t0 < r.arrival:
                pass
            cap = int(self.stats[PR.table_of(r.ops)].rows.sum())
            need = sum((max(cap, 1) * w + 255) & ~255 for w in _widths(r.ops))
            base = self.arena.alloc(need)
            if base is None:
                raise VirginianError("a request's output buffers do not fit in the arena")
            outs, off = [], base
            for w in _widths(r.ops):
                outs.append(self.arena.base + off)
                off += (max(cap, 1) * w + 255) & ~255
            q = _Query(self.db, r, max(cap, 1), outs)
            n = C.c_ulonglong(0)
            if lib.vg_query_run(q.h, C.byref(n)) != VIRG_SUCCESS:
                raise VirginianError("vg_query_run failed")
            q.last_rows = n.value
            if self.keep:
                r.result = q.fetch()
            r.ready = time.perf_counter() - t0
            r.rows = n.value
            q.close()
            self.arena.release(base, need)
            self.peak_bytes = max(self.peak_bytes, need)
            self.reserved += need

    def run(self, reqs: list[Request]) -> dict:
        """Serve the requests at their arrival times; returns a summary."""
        if self.mode == "plain":
            self._run_plain(reqs)
            end = max(r.ready for r in reqs) if reqs else 0.0
            lat = np.array([r.ready - r.arrival for r in reqs]) * 1e3
            return {"requests": len(reqs), "seconds": end, "queries_per_s": len(reqs) / end if end else 0.0,
                    "p50_ms": float(np.percentile(lat, 50)) if len(lat) else 0.0,
                    "p99_ms": float(np.percentile(lat, 99)) if len(lat) else 0.0,
                    "reruns": 0, "peak_arena_bytes": self.peak_bytes, "reserved_bytes": self.reserved,
                    "host_seconds": {}}
        pending = deque(sorted(reqs, key=lambda r: r.arrival))
        waiting: deque = deque()
        reruns: deque = deque()
        free = list(self.runners)
        rerun_free = [self.rerun_runner]
        inflight = []
        done = 0
        # admission is tried again only when something changed: an arrival,
        # a finished batch (arena and a runner freed) or a batch launched
        # (the waiting line moved); otherwise it would fail the same way
        changed = True
        t0 = time.perf_counter()
        while done < len(reqs):
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
                changed = True
            progressed = False
            # reruns first (they arrived earlier), then new requests
            prof = self.prof
            admit_now, changed = changed, False
            for queue, pool in ((reruns, rerun_free), (waiting, free)):
                if admit_now and queue and pool:
                    ta = time.perf_counter()
                    batch, block = self._admit(queue)
                    tb = time.perf_counter()
                    prof["admit"] += tb - ta
                    if batch:
                        runner = pool.pop()
                        runner.reload([q for _, q in batch])
                        runner.launch()
                        inflight.append((runner, batch, block, pool))
                        progressed = changed = True
                        prof["launch"] += time.perf_counter() - tb
                        prof["batches"] += 1
                    elif not inflight:
                        raise VirginianError("a request's output buffers do not fit in the arena")
            still = []
            for runner, batch, block, pool in inflight:
                if runner.done():
                    ta = time.perf_counter()
                    done += self._finish(runner, batch, block, t0, reruns)
                    prof["collect"] += time.perf_counter() - ta
                    pool.append(runner)
                    progressed = changed = True
                else:
                    still.append((runner, batch, block, pool))
            inflight = still
            if not progressed:
                ta = time.perf_counter()
                time.sleep(0)
                prof["idle"] += time.perf_counter() - ta
        end = max(r.ready for r in reqs) if reqs else 0.0
        lat = np.array([r.ready - r.arrival for r in reqs]) * 1e3
        return {"requests": len(reqs), "seconds": end, "queries_per_s": len(reqs) / end if end else 0.0,
                "p50_ms": float(np.percentile(lat, 50)) if len(lat) else 0.0,
                "p99_ms": float(np.percentile(lat, 99)) if len(lat) else 0.0,
                "reruns": sum(r.rerun for r in reqs), "peak_arena_bytes": self.peak_bytes,
                "reserved_bytes": self.reserved,
                "host_seconds": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in self.prof.items()}}
