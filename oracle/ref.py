"""ORACLE -- test infrastructure only.

ctypes driver for oracle/_ref/libvirgref.so: the reference's own CPU VM
compiled from its sources by oracle/build_ref.sh (see ref_harness.c).  Used
by tests/ to check the C restatement (virg_oracle.c) against the code it
restates.  The library keeps one reference database per process.
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile
from pathlib import Path

import numpy as np

import virginian_amd as va

PATH = Path(__file__).resolve().parent / "_ref" / "libvirgref.so"
_lib = None


def available() -> bool:
    return PATH.exists()


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = C.CDLL(str(PATH))
        _lib.ref_open.argtypes = [C.c_char_p, C.c_int, C.c_int]
        _lib.ref_open.restype = C.c_int
        _lib.ref_open_file.argtypes = [C.c_char_p, C.c_int, C.c_int]
        _lib.ref_open_file.restype = C.c_int
        _lib.ref_db_close.argtypes = []
        _lib.ref_db_close.restype = C.c_int
        _lib.ref_close.argtypes = []
        _lib.ref_table.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_char_p),
                                   C.POINTER(C.c_int)]
        _lib.ref_table.restype = C.c_int
        _lib.ref_insert.argtypes = [C.c_uint, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_int), C.c_ulonglong]
        _lib.ref_insert.restype = C.c_int
        _lib.ref_run.argtypes = [C.POINTER(va.Op), C.c_uint, C.POINTER(C.c_void_p), C.c_uint,
                                 C.c_ulonglong]
        _lib.ref_run.restype = C.c_longlong
        _lib.ref_run_reader.argtypes = [C.POINTER(va.Op), C.c_uint, C.c_void_p, C.c_uint, C.c_ulonglong,
                                        C.POINTER(C.c_ulonglong)]
        _lib.ref_run_reader.restype = C.c_longlong
    return _lib


class RefDB:
    """One reference database (virg_init .. virg_close minus CUDA)."""

    def __init__(self, multi: bool = False, threads: int = 4, path: str | None = None,
                 existing: bool = False):
        """A new database (in a temporary file unless `path` is given), or
        with existing=True the database file at `path`."""
        self.dir = None
        if path is None:
            self.dir = tempfile.TemporaryDirectory(prefix="virgref")
            path = os.path.join(self.dir.name, "ref.db")
        self.path = str(path)
        opener = lib().ref_open_file if existing else lib().ref_open
        if opener(self.path.encode(), 1 if multi else 0, threads) != 1:
            raise RuntimeError(f"reference could not open {self.path}")
        self.rows = 0

    def close(self) -> None:
        """Free the database without writing the file."""
        lib().ref_close()
        if self.dir is not None:
            self.dir.cleanup()

    def write_close(self) -> None:
        """Write the database file (the reference's virg_db_close) and free."""
        ok = lib().ref_db_close()
        if self.dir is not None:
            self.dir.cleanup()
        if ok != 1:
            raise RuntimeError("reference virg_db_close failed")

    def table(self, name: str, key: np.ndarray, cols: list[tuple[str, np.ndarray]]) -> int:
        """Create a table and insert every row through virg_table_insert."""
        kt = {np.dtype(np.int32): va.VIRG_INT, np.dtype(np.int64): va.VIRG_INT64}[key.dtype]
        tmap = {np.dtype(np.int32): va.VIRG_INT, np.dtype(np.int64): va.VIRG_INT64,
                np.dtype(np.float32): va.VIRG_FLOAT, np.dtype(np.float64): va.VIRG_DOUBLE,
                np.dtype(np.int8): va.VIRG_CHAR}
        names = (C.c_char_p * len(cols))(*[n.encode() for n, _ in cols])
        types = (C.c_int * len(cols))(*[tmap[c.dtype] for _, c in cols])
        tid = lib().ref_table(name.encode(), kt, len(cols), names, types)
        if tid < 0:
            raise RuntimeError("ref_table failed")
        arrs = [np.ascontiguousarray(c) for _, c in cols]
        key = np.ascontiguousarray(key)
        ptrs = (C.c_void_p * len(arrs))(*[x.ctypes.data for x in arrs])
        widths = (C.c_int * len(arrs))(*[x.dtype.itemsize for x in arrs])
        if lib().ref_insert(tid, key.ctypes.data, key.dtype.itemsize, len(arrs), ptrs, widths,
                            len(key)) != 1:
            raise RuntimeError("ref_insert failed")
        self.rows = len(key)
        return tid

    def run(self, ops, names: list[str], types: list[int], cap: int) -> dict:
        """Run an op program (Op array, n) and return the logical result."""
        arr, n = ops
        outs = [np.empty(cap, dtype=va.NP_TYPE[t]) for t in types]
        ptrs = (C.c_void_p * max(len(outs), 1))(*[o.ctypes.data for o in outs])
        rows = lib().ref_run(arr, n, ptrs, len(outs), cap)
        if rows < 0:
            raise RuntimeError("ref_run failed")
        return {nm: o[:rows] for nm, o in zip(names, outs)}

    def run_reader(self, ops, row_bytes: int, cap: int) -> tuple[np.ndarray, int]:
        """Run an op program, then read its result with the reference's own
        virg_reader_row: the reader's buffer (row_bytes of it) after every
        call up to and including the first VIRG_FAIL, and the number of
        calls that returned VIRG_SUCCESS."""
        arr, n = ops
        out = np.zeros((cap, row_bytes), dtype=np.uint8)
        ok = C.c_ulonglong(0)
        calls = lib().ref_run_reader(arr, n, out.ctypes.data, row_bytes, cap, C.byref(ok))
        if calls < 0:
            raise RuntimeError("ref_run_reader failed")
        return out[:calls], ok.value
