"""Serving loop: queries arrive over time and are run with continuous
batching against resident tables.

The loop itself is the host library's (vg_server_* in include/virg_gpu.h,
virginian_amd/csrc/gpu/serve.hip): one host thread admits arrived requests
in arrival order while their result buffers fit in an HBM arena
(admission by footprint), runs one admitted request alone (its own count
and write launches) or several as one persistent batch on a free runner
stream while earlier work still runs on the others, collects finished
runs from pinned host memory, and reruns at its exact size a result that
did not fit.  This module is a thin driver: it submits requests with
their arrival times, runs the loop and reads each request's outcome.

Latency is measured from a request's arrival to its result complete in
HBM (with keep_results, in host memory).  No counterpart in the reference
(its queries run one at a time, src/vm/execute.c:28-102).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import NP_TYPE, OP_ResultColumn, VIRG_SUCCESS, VirginianError, lib, make_ops
from . import predict as PR

# vg_server_config.sizing (include/virg_gpu.h)
SIZE_ALL, SIZE_ESTIMATE, SIZE_NET = 0, 1, 2
# vg_server_req.state
ST_DONE, ST_FAILED = 3, 4


class ServerConfig(C.Structure):
    """vg_server_config (include/virg_gpu.h)."""
    _fields_ = [("arena_bytes", C.c_ulonglong), ("runners", C.c_uint), ("max_batch", C.c_uint),
                ("tiles_per_item", C.c_uint), ("sizing", C.c_uint), ("margin", C.c_float),
                ("single", C.c_uint), ("split", C.c_uint), ("grid_div", C.c_uint), ("keep", C.c_uint),
                ("net", C.c_void_p), ("batch_line", C.c_uint), ("costly", C.c_ulonglong),
                ("host_arena", C.c_uint)]


class ServerStats(C.Structure):
    """vg_server_stats (include/virg_gpu.h)."""
    _fields_ = [("done", C.c_ulonglong), ("failed", C.c_ulonglong), ("reruns", C.c_ulonglong),
                ("lone", C.c_ulonglong), ("batches", C.c_ulonglong), ("reserved_bytes", C.c_ulonglong),
                ("arena_peak", C.c_ulonglong), ("in_flight_peak", C.c_ulonglong),
                ("admit_s", C.c_double), ("launch_s", C.c_double), ("collect_s", C.c_double)]


class ServerReq(C.Structure):
    """vg_server_req (include/virg_gpu.h)."""
    _fields_ = [("state", C.c_uint), ("reruns", C.c_uint), ("rows", C.c_ulonglong),
                ("capacity", C.c_ulonglong), ("arrival", C.c_double), ("ready", C.c_double)]


_BOUND = False


def _bind() -> None:
    global _BOUND
    if _BOUND:
        return
    P = C.c_void_p
    sig = {
        "vg_server_create": (P, [P, C.POINTER(ServerConfig), C.c_char_p, C.c_size_t]),
        "vg_server_free": (None, [P]),
        "vg_server_table": (C.c_int, [P, C.c_uint, P]),
        "vg_server_submit": (C.c_int, [P, P, C.c_uint, C.c_double, C.POINTER(C.c_ulonglong)]),
        "vg_server_step": (C.c_int, [P]),
        "vg_server_run": (C.c_int, [P, C.c_double]),
        "vg_server_now": (C.c_double, [P]),
        "vg_server_request": (C.c_int, [P, C.c_ulonglong, C.POINTER(ServerReq)]),
        "vg_server_fetch": (C.c_int, [P, C.c_ulonglong, C.POINTER(P), C.c_ulonglong]),
        "vg_server_release": (C.c_int, [P, C.c_ulonglong]),
        "vg_server_get_stats": (C.c_int, [P, C.POINTER(ServerStats)]),
        "vg_server_error": (C.c_char_p, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _BOUND = True


@dataclass
class Request:
    arrival: float                  # seconds after the start
    ops: list
    rid: int = 0
    ready: float = -1.0             # seconds after the start
    rows: int = 0
    rerun: bool = False
    cap: int = -1                   # output rows reserved for its last run
    result: dict = field(default_factory=dict)


# bytes per value of each virg_t (INT, INT64, FLOAT, DOUBLE, CHAR)
_WIDTH = [4, 8, 4, 8, 1]


class Server:
    """Modes (how result buffers are sized, and how requests run):
    "sized": by the result-size network (or, without one, the independence
    estimate), requests batched, two-pass; "estimate": by the independence
    estimate; "single": as "sized", every run a single-pass batch;
    "two-pass": buffers for every row of the table (the method without a
    predictor); "plain": no batching -- each request alone, in arrival
    order, run with its own launch pair and waited for, then the next."""

    def __init__(self, db, table_ids: list[int], arena_bytes: int, mode: str = "single", predictor=None,
                 margin: float = 0.05, runners: int = 3, max_batch: int = 64, tiles_per_item: int = 8,
                 keep_results: bool = True, split: int = 0, grid_div: int = 0, batch_line: int = 0,
                 costly: int = 0, host_arena: bool = False):
        _bind()
        if mode not in ("sized", "estimate", "single", "two-pass", "plain"):
            raise ValueError(f"unknown serving mode {mode}")
        self.db = db
        self.mode = mode
        self.keep = keep_results
        sizing = SIZE_ALL
        if mode in ("sized", "single", "estimate"):
            sizing = SIZE_NET if predictor is not None and mode != "estimate" else SIZE_ESTIMATE
        if mode == "plain":
            runners, max_batch = 1, 1
        cfg = ServerConfig(arena_bytes, runners, max_batch, tiles_per_item, sizing, margin,
                           1 if mode == "single" else 0, split, grid_div, 1 if keep_results else 0,
                           predictor.h if sizing == SIZE_NET else None, batch_line, costly,
                           1 if host_arena else 0)
        err = C.create_string_buffer(256)
        self.h = lib.vg_server_create(C.cast(db.ptr, C.c_void_p), C.byref(cfg), err, 256)
        if not self.h:
            raise VirginianError(err.value.decode() or "vg_server_create failed")
        if sizing != SIZE_ALL:
            for t in table_ids:
                st = PR.TableStats(db, t)
                if lib.vg_server_table(self.h, t, C.addressof(st.c_stats())) != VIRG_SUCCESS:
                    raise VirginianError("vg_server_table failed")

    def close(self) -> None:
        if self.h:
            lib.vg_server_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _error(self, what: str) -> VirginianError:
        msg = lib.vg_server_error(self.h)
        return VirginianError(f"{what}: {msg.decode() if msg else ''}")

    def stats(self) -> dict:
        st = ServerStats()
        lib.vg_server_get_stats(self.h, C.byref(st))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def run(self, reqs: list[Request], timeout_s: float = 600.0) -> dict:
        """Serve the requests at their arrival times; returns a summary."""
        ids = []
        for r in reqs:
            arr, n, keep = make_ops(r.ops)
            i = C.c_ulonglong(0)
            if lib.vg_server_submit(self.h, C.cast(arr, C.c_void_p), n, r.arrival, C.byref(i)) != VIRG_SUCCESS:
                raise self._error("vg_server_submit failed")
            ids.append(i.value)
        if lib.vg_server_run(self.h, timeout_s) != VIRG_SUCCESS:
            raise self._error("vg_server_run failed")
        for r, i in zip(reqs, ids):
            q = ServerReq()
            lib.vg_server_request(self.h, i, C.byref(q))
            if q.state != ST_DONE:
                raise self._error(f"request {r.rid} failed")
            r.ready, r.rows, r.rerun, r.cap = q.ready, q.rows, q.reruns > 0, q.capacity
            if self.keep:
                types = [o[1] for o in r.ops if o[0] == OP_ResultColumn]
                names = [o[4] if isinstance(o[4], str) else f"c{k}"
                         for k, o in enumerate([o for o in r.ops if o[0] == OP_ResultColumn])]
                outs = [np.empty(r.rows, dtype=NP_TYPE[t]) for t in types]
                ptrs = (C.c_void_p * max(len(outs), 1))(*[o.ctypes.data for o in outs])
                if lib.vg_server_fetch(self.h, i, ptrs, r.rows) != VIRG_SUCCESS:
                    raise self._error("vg_server_fetch failed")
                lib.vg_server_release(self.h, i)
                r.result = dict(zip(names, outs))
        st = self.stats()
        end = max(r.ready for r in reqs) if reqs else 0.0
        lat = np.array([r.ready - r.arrival for r in reqs]) * 1e3
        return {"requests": len(reqs), "seconds": end, "queries_per_s": len(reqs) / end if end else 0.0,
                "p50_ms": float(np.percentile(lat, 50)) if len(lat) else 0.0,
                "p90_ms": float(np.percentile(lat, 90)) if len(lat) else 0.0,
                "p99_ms": float(np.percentile(lat, 99)) if len(lat) else 0.0,
                "reruns": sum(r.rerun for r in reqs), "peak_arena_bytes": st["arena_peak"],
                "reserved_bytes": st["reserved_bytes"],
                "host_seconds": {"admit": round(st["admit_s"], 4), "launch": round(st["launch_s"], 4),
                                 "collect": round(st["collect_s"], 4), "lone": st["lone"],
                                 "batches": st["batches"]}}
