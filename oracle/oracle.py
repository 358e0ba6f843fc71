"""ORACLE loader -- test infrastructure only.

ctypes binding of oracle/liboracle.so (the C restatement of the reference
CPU VM in virg_oracle.c).  Imported only by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg, always as the checker.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import virginian_amd as va

_PATH = Path(__file__).resolve().parent / "liboracle.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not _PATH.exists():
            raise RuntimeError(f"{_PATH} missing: run python -m virginian_amd.build")
        _lib = C.CDLL(str(_PATH))
        common = [C.POINTER(va.Op), C.c_uint, C.c_uint, C.POINTER(C.POINTER(va.TabletMeta)),
                  C.c_uint, C.POINTER(C.c_void_p), C.c_uint, C.c_ulonglong]
        _lib.oracle_select.restype = C.c_longlong
        _lib.oracle_select.argtypes = common + [C.POINTER(C.c_int), C.c_char_p, C.c_size_t]
        _lib.oracle_select_mt.restype = C.c_longlong
        _lib.oracle_select_mt.argtypes = common + [C.c_int, C.c_int, C.POINTER(C.c_int),
                                                   C.c_char_p, C.c_size_t]
    return _lib


class OracleError(RuntimeError):
    pass


def _ops(db: va.Virginian, prog):
    """SQL string, virg_vm pointer or op tuples -> (Op array, n, names, types, keep)."""
    if isinstance(prog, str):
        vm = db.compile(prog)
        try:
            arr, n = va.ops_array(vm)
            names, types = va.result_schema(vm)
            # ResultColumn names live in the vm; the oracle never reads them
            return arr, n, names, types, None
        finally:
            db.free_vm(vm)
    if isinstance(prog, list):
        arr, n, keep = va.make_ops(prog)
        names = [o[4] for o in prog if o[0] == va.OP_ResultColumn]
        types = [o[1] for o in prog if o[0] == va.OP_ResultColumn]
        return arr, n, names, types, keep
    arr, n = va.ops_array(prog)
    names, types = va.result_schema(prog)
    return arr, n, names, types, None


def select(db: va.Virginian, prog, table: str | None = None, threads: int = 0,
           ordered: bool = True, count_only: bool = False,
           max_tablets: int | None = None, tablets: tuple[int, int] | None = None) -> tuple[dict, int]:
    """Run the program's data-parallel section on the CPU oracle over the
    table named by its Table op (or `table`).  Returns (columns, illformed).
    threads = 0 runs the single-threaded reference-order VM; tablets =
    (i, j) runs it over the table's tablets i..j-1 only."""
    arr, n, names, types, _keep = _ops(db, prog)
    tid = None
    par = None
    for i in range(n):
        if arr[i].op == va.OP_Table and tid is None:
            tid = arr[i].p1
        if arr[i].op == va.OP_Parallel and par is None:
            par = i
    if table is not None:
        tid = db.table_id(table)
    if tid is None or par is None:
        raise OracleError("program has no Table/Parallel op")
    tabs = db.tablets(tid)
    if max_tablets is not None:
        tabs = tabs[:max_tablets]
    if tablets is not None:
        tabs = tabs[tablets[0]:tablets[1]]
    tarr = (C.POINTER(va.TabletMeta) * len(tabs))(*tabs)
    cap = sum(t.contents.rows for t in tabs)
    outs = [] if count_only else [np.empty(cap, dtype=va.NP_TYPE[t]) for t in types]
    ptrs = (C.c_void_p * max(len(outs), 1))(*[o.ctypes.data for o in outs]) if outs else None
    ill = C.c_int(0)
    err = C.create_string_buffer(200)
    if threads:
        rows = lib().oracle_select_mt(arr, n, par + 1, tarr, len(tabs), ptrs, len(outs), cap,
                                      threads, 1 if ordered else 0, C.byref(ill), err, 200)
    else:
        rows = lib().oracle_select(arr, n, par + 1, tarr, len(tabs), ptrs, len(outs), cap,
                                   C.byref(ill), err, 200)
    if rows < 0:
        raise OracleError(err.value.decode())
    cols = {} if count_only else {nm: o[:rows] for nm, o in zip(names, outs)}
    if count_only:
        cols = {"__rows__": rows}
    return cols, ill.value
