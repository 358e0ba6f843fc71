"""virginian_amd -- MI355X-native opcode-VM query executor.

Python bindings over the C ABI of virginian_amd/lib/libvirginian_amd.so
(declared in include/virginian.h and include/virg_gpu.h).  The library is
built in-tree by ``python -m virginian_amd.build`` (or
``__graft_entry__.build()``); importing this package without it raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# VIRGINIAN_AMD_LIB selects another build of the same library (A/B timing)
_LIB_PATH = Path(os.environ.get("VIRGINIAN_AMD_LIB") or
                 Path(__file__).resolve().parent / "lib" / "libvirginian_amd.so")

VIRG_FAIL, VIRG_SUCCESS = 0, 1
VG_QUERY_LIGHT, VG_QUERY_NO_OUTPUT, VG_QUERY_INTERP = 1, 2, 4
VIRG_MAX_COLUMNS = 16
VIRG_MAX_COLUMN_NAME = 16
VIRG_MAX_TABLES = 16
VIRG_MAX_TABLE_NAME = 32
VIRG_MEM_TABLETS = 64
VIRG_OPS = 32
VIRG_REGS = 16
VIRG_GLOBAL_REGS = 16
VIRG_ROW_BUFFER = 256
VIRG_TABLET_SIZE = 8 * 1024 * 1024

# opcodes (reference: src/virginian.h:166-193)
(OP_Table, OP_ResultColumn, OP_Parallel, OP_Finish, OP_Column, OP_Rowid, OP_Result,
 OP_Converge, OP_Invalid, OP_Cast, OP_Integer, OP_Float, OP_Le, OP_Lt, OP_Ge, OP_Gt,
 OP_Eq, OP_Neq, OP_Add, OP_Sub, OP_Mul, OP_Div, OP_And, OP_Or, OP_Not, OP_Nop) = range(26)

# value types (reference: src/virginian.h:197-205)
VIRG_INT, VIRG_INT64, VIRG_FLOAT, VIRG_DOUBLE, VIRG_CHAR, VIRG_STRING, VIRG_NULL = range(7)
NP_TYPE = {VIRG_INT: np.int32, VIRG_INT64: np.int64, VIRG_FLOAT: np.float32,
           VIRG_DOUBLE: np.float64, VIRG_CHAR: np.int8}


class TabletInfo(C.Structure):
    _fields_ = [("id", C.c_uint), ("used", C.c_int), ("disk_slot", C.c_uint)]


class TabletMeta(C.Structure):
    """virg_tablet_meta (reference: src/virginian.h:251-302), 704 bytes."""
    _fields_ = [
        ("rows", C.c_uint), ("key_type", C.c_int), ("key_stride", C.c_size_t),
        ("key_pointer_stride", C.c_size_t), ("id", C.c_uint), ("next", C.c_uint),
        ("last_tablet", C.c_int), ("in_table", C.c_int), ("table_id", C.c_uint),
        ("key_block", C.c_size_t), ("key_pointers_block", C.c_size_t),
        ("fixed_block", C.c_size_t), ("variable_block", C.c_size_t), ("size", C.c_size_t),
        ("possible_rows", C.c_uint), ("row_stride", C.c_size_t), ("fixed_columns", C.c_uint),
        ("fixed_name", (C.c_char * VIRG_MAX_COLUMN_NAME) * VIRG_MAX_COLUMNS),
        ("fixed_type", C.c_int * VIRG_MAX_COLUMNS),
        ("fixed_stride", C.c_size_t * VIRG_MAX_COLUMNS),
        ("fixed_offset", C.c_size_t * VIRG_MAX_COLUMNS),
        ("info", C.POINTER(TabletInfo)),
        ("_pad", C.c_char * 8),
    ]


class Var(C.Union):
    _fields_ = [("i", C.c_int), ("f", C.c_float), ("li", C.c_longlong), ("d", C.c_double),
                ("c", C.c_char), ("s", C.c_char_p)]


class Op(C.Structure):
    """virg_op (reference: src/virginian.h:326-332), 24 bytes."""
    _fields_ = [("op", C.c_int), ("p1", C.c_int), ("p2", C.c_int), ("p3", C.c_int), ("p4", Var)]


class ResultNode(C.Structure):
    pass


ResultNode._fields_ = [("id", C.c_uint), ("next", C.POINTER(ResultNode))]


class VM(C.Structure):
    """virg_vm (reference: src/virginian.h:356-377), 1016 bytes."""
    _fields_ = [
        ("pc", C.c_uint), ("stmt", Op * VIRG_OPS), ("num_ops", C.c_uint),
        ("global_reg", Var * VIRG_GLOBAL_REGS), ("type", C.c_int * VIRG_GLOBAL_REGS),
        ("table", C.c_uint * 1), ("num_tables", C.c_uint),
        ("head_result", C.POINTER(ResultNode)), ("tail_result", C.POINTER(ResultNode)),
        ("timing1", C.c_float), ("timing2", C.c_float), ("timing3", C.c_float),
    ]


class DB(C.Structure):
    _fields_ = [
        ("num_tablets", C.c_uint), ("alloced_tablets", C.c_uint), ("tablet_id_counter", C.c_uint),
        ("block_size", C.c_size_t),
        ("tables", (C.c_char * VIRG_MAX_TABLE_NAME) * VIRG_MAX_TABLES),
        ("first_tablet", C.c_uint * VIRG_MAX_TABLES), ("last_tablet", C.c_uint * VIRG_MAX_TABLES),
        ("write_cursor", C.c_uint * VIRG_MAX_TABLES), ("table_tablets", C.c_uint * VIRG_MAX_TABLES),
        ("table_status", C.c_int * VIRG_MAX_TABLES), ("tablet_info", C.POINTER(TabletInfo)),
    ]


class State(C.Structure):
    """virginian (reference: src/virginian.h:473-504), 1976 bytes."""
    _fields_ = [
        ("db", DB), ("tablet_slot_ids", C.c_uint * VIRG_MEM_TABLETS),
        ("tablet_slot_status", C.c_int * VIRG_MEM_TABLETS),
        ("tablet_slots_taken", C.c_uint), ("tablet_slot_counter", C.c_uint),
        ("tablet_slots", C.POINTER(TabletMeta) * VIRG_MEM_TABLETS), ("gpu_slots", C.c_void_p),
        ("slot_lock", C.c_byte * 40), ("dbfd", C.c_int), ("threads_per_block", C.c_uint),
        ("multi_threads", C.c_uint), ("use_multi", C.c_int), ("use_gpu", C.c_int),
        ("use_stream", C.c_int), ("use_mmap", C.c_int),
    ]


class Reader(C.Structure):
    _fields_ = [("vm", C.POINTER(VM)), ("res", C.POINTER(TabletMeta)), ("row", C.c_uint),
                ("buffer", C.c_char * VIRG_ROW_BUFFER)]


class VirginianError(RuntimeError):
    pass


class StreamStats(C.Structure):
    """vg_stream_stats (include/virg_gpu.h)."""
    _fields_ = [("rows_scanned", C.c_ulonglong), ("rows_out", C.c_ulonglong),
                ("h2d_bytes", C.c_ulonglong), ("d2h_bytes", C.c_ulonglong),
                ("seconds", C.c_double), ("batches", C.c_uint)]


class PBatchConfig(C.Structure):
    """vg_pbatch_config (include/virg_gpu.h)."""
    _fields_ = [("tiles_per_item", C.c_uint), ("grid_div", C.c_uint), ("split", C.c_uint),
                ("order", C.c_uint), ("direct", C.c_uint), ("db", C.c_void_p), ("single", C.c_uint),
                ("jit", C.c_uint)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "db"}


def _load() -> C.CDLL:
    if not _LIB_PATH.exists():
        raise VirginianError(
            f"{_LIB_PATH} is missing: build it with `python -m virginian_amd.build` "
            "(the executor has no CPU fallback)")
    lib = C.CDLL(str(_LIB_PATH), mode=C.RTLD_GLOBAL)
    P, V = C.c_void_p, C.POINTER(State)
    sig = {
        "virg_init": (C.c_int, [V]), "virg_close": (C.c_int, [V]),
        "virg_db_create": (C.c_int, [V, C.c_char_p]), "virg_db_open": (C.c_int, [V, C.c_char_p]),
        "virg_db_close": (C.c_int, [V]), "virg_db_clear": (C.c_int, [V, C.c_uint]),
        "virg_db_write": (C.c_int, [V, C.c_uint]),
        "virg_db_load": (C.c_int, [V, C.c_uint, C.POINTER(C.POINTER(TabletMeta))]),
        "virg_tablet_lock": (C.c_int, [V, C.c_uint]),
        "virg_tablet_unlock": (C.c_int, [V, C.c_uint]),
        "virg_gpu_getsizes": (C.POINTER(C.c_size_t), []),
        "virg_table_create": (C.c_int, [V, C.c_char_p, C.c_int]),
        "virg_table_addcolumn": (C.c_int, [V, C.c_uint, C.c_char_p, C.c_int]),
        "virg_table_insert": (C.c_int, [V, C.c_uint, P, P, P]),
        "virg_table_getid": (C.c_int, [V, C.c_char_p, C.POINTER(C.c_uint)]),
        "virg_table_numrows": (C.c_int, [V, C.c_uint, C.POINTER(C.c_uint)]),
        "virg_table_appendtablet": (C.c_int, [V, C.c_uint, C.c_uint, C.c_uint, P, C.POINTER(P)]),
        "virg_table_tablets": (C.c_int, [V, C.c_uint, C.POINTER(C.POINTER(TabletMeta)), C.c_int]),
        "virg_gen_table": (C.c_int, [V, C.c_char_p, C.c_uint64, C.c_uint, C.c_uint32, C.c_int]),
        "virg_gen_rows_mt": (None, [C.c_uint32, C.c_uint64, C.c_uint64, P, P, P, P, P, P, P, C.c_int]),
        "virg_philox": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "virg_log": (C.c_double, [C.c_double]),
        "virg_gen_normal": (C.c_double, [C.c_uint32, C.c_uint32, C.c_uint64]),
        "virg_gen_uniform": (C.c_double, [C.c_uint32, C.c_uint32, C.c_uint64]),
        "virg_vm_init": (C.POINTER(VM), []),
        "virg_vm_cleanup": (None, [V, C.POINTER(VM)]),
        "virg_vm_addop": (C.c_int, [C.POINTER(VM), C.c_int, C.c_int, C.c_int, C.c_int, Var]),
        "virg_vm_execute": (C.c_int, [V, C.POINTER(VM)]),
        "virg_vm_freeresults": (C.c_int, [V, C.POINTER(VM)]),
        "virg_sql": (C.c_int, [V, C.c_char_p, C.POINTER(VM)]),
        "virg_sql_compile": (C.c_int, [V, C.c_char_p, C.POINTER(VM), C.c_char_p, C.c_size_t]),
        "virg_query": (C.c_int, [V, C.POINTER(C.POINTER(Reader)), C.c_char_p]),
        "virg_release": (None, [V, C.POINTER(Reader)]),
        "virg_reader_row": (C.c_int, [V, C.POINTER(Reader)]),
        "virg_reader_getrows": (C.c_int, [V, C.POINTER(Reader), C.POINTER(C.c_uint)]),
        "virg_result_columns": (C.c_int, [V, C.POINTER(VM), C.POINTER(P), C.POINTER(C.c_ulonglong)]),
        "virg_opstring": (C.c_char_p, [C.c_int]),
        "virg_tablet_check": (C.c_int, [C.POINTER(TabletMeta)]),
        "virg_vm_gpu": (C.c_int, [V, C.POINTER(VM), C.POINTER(C.POINTER(TabletMeta)),
                                  C.POINTER(C.POINTER(TabletMeta)), C.c_uint]),
        "vg_query_create": (P, [V, C.POINTER(Op), C.c_uint, C.c_ulonglong, C.c_int, C.c_char_p, C.c_size_t]),
        "vg_query_free": (None, [P]),
        "vg_query_launch": (C.c_int, [P]),
        "vg_query_sync": (C.c_int, [P, C.POINTER(C.c_ulonglong)]),
        "vg_query_run": (C.c_int, [P, C.POINTER(C.c_ulonglong)]),
        "vg_query_fetch": (C.c_int, [P, C.POINTER(P), C.c_ulonglong]),
        "vg_query_time": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_ulonglong)]),
        "vg_query_time_passes": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_double),
                                           C.POINTER(C.c_double), C.POINTER(C.c_ulonglong)]),
        "vg_query_info": (C.c_int, [P, C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                    C.POINTER(C.c_uint), C.POINTER(C.c_uint)]),
        "vg_device_info": (C.c_int, [V, C.c_char_p, C.c_size_t, C.POINTER(C.c_int)]),
        "vg_gen_table": (C.c_int, [V, C.c_char_p, C.c_ulonglong, C.c_uint, C.c_uint]),
        "vg_stream_select": (C.c_int, [V, C.POINTER(Op), C.c_uint, C.POINTER(P), C.c_ulonglong,
                                       C.c_ulonglong, C.POINTER(StreamStats), C.c_char_p, C.c_size_t]),
        "vg_batch_create": (P, [C.POINTER(P), C.c_uint, C.c_char_p, C.c_size_t]),
        "vg_batch_free": (None, [P]),
        "vg_batch_launch": (C.c_int, [P]),
        "vg_batch_sync": (C.c_int, [P, C.POINTER(C.c_ulonglong)]),
        "vg_batch_time": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "vg_host_memory": (C.c_int, [V, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
        "vg_host_pin_limit": (C.c_int, [V, C.c_ulonglong]),
        "vg_tablet_limit": (C.c_uint, [V, C.c_uint]),
        "vg_pager_stats": (C.c_int, [V, C.POINTER(C.c_uint), C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong),
                                     C.c_uint, C.POINTER(C.c_int)]),
        "vg_pbatch_create": (P, [C.POINTER(P), C.c_uint, C.POINTER(PBatchConfig), C.c_char_p, C.c_size_t]),
        "vg_pbatch_free": (None, [P]),
        "vg_pbatch_configure": (C.c_int, [P, C.POINTER(PBatchConfig), C.c_char_p, C.c_size_t]),
        "vg_pbatch_get_config": (C.c_int, [P, C.POINTER(PBatchConfig), C.POINTER(C.c_uint)]),
        "vg_pbatch_launch": (C.c_int, [P]),
        "vg_pbatch_sync": (C.c_int, [P, C.POINTER(C.c_ulonglong), C.POINTER(C.c_double)]),
        "vg_pbatch_time": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "vg_pbatch_results": (C.c_int, [P, C.POINTER(C.c_ulonglong), C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
        "vg_pbatch_reload": (C.c_int, [P, C.POINTER(P), C.c_uint, C.c_char_p, C.c_size_t]),
        "vg_pbatch_done": (C.c_int, [P]),
        "vg_query_set_outputs": (C.c_int, [P, C.POINTER(P), C.c_ulonglong]),
        "vg_query_out_widths": (C.c_uint, [P, C.POINTER(C.c_uint)]),
        "vg_device_alloc": (P, [V, C.c_size_t]),
        "vg_device_free": (None, [V, P]),
        "vg_pbatch_tune": (C.c_int, [P, C.c_int, C.POINTER(PBatchConfig), C.POINTER(C.c_double),
                                     C.c_char_p, C.c_size_t]),
        "vg_plain_time": (C.c_int, [C.POINTER(P), C.c_uint, C.c_int, C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
        "vg_host_alloc": (P, [C.c_size_t]),
        "vg_host_free": (None, [P]),
        "vg_link_probe": (C.c_int, [V, C.c_size_t, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "vg_decode_program": (C.c_int, [C.POINTER(Op), C.c_uint, C.c_uint, C.POINTER(TabletMeta),
                                        P, C.c_size_t, C.c_char_p, C.c_size_t]),
        "vg_jit_policy": (C.c_int, [C.c_int]),
        "vg_select_noalloc": (C.c_int, [C.POINTER(Op), C.c_uint, C.POINTER(P), C.POINTER(C.c_int), C.c_uint, P,
                                        C.c_int, C.c_ulonglong, C.POINTER(P), C.c_ulonglong, P, P, C.c_char_p,
                                        C.c_size_t]),
        "vg_query_kernel": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_uint),
                                      C.POINTER(C.c_uint), C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "vg_jit_source": (C.c_long, [C.POINTER(Op), C.c_uint, C.c_uint, C.POINTER(TabletMeta), C.c_char_p,
                                     C.c_size_t, C.c_char_p, C.c_size_t]),
        "vg_jit_compile_check": (C.c_int, [C.POINTER(Op), C.c_uint, C.c_uint, C.POINTER(TabletMeta),
                                           C.POINTER(C.c_double), C.c_char_p, C.c_size_t]),
        "vg_jitb_check": (C.c_uint, [C.POINTER(C.POINTER(Op)), C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                     C.POINTER(C.POINTER(TabletMeta)), C.POINTER(C.c_int), C.c_uint,
                                     C.POINTER(C.c_uint32), C.c_int, C.POINTER(C.c_double), C.c_char_p,
                                     C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


class _Lib:
    """The native library, loaded on first use (so that the package can be
    imported to build it); any use without it raises VirginianError."""

    _cdll = None

    def __getattr__(self, name: str):
        if _Lib._cdll is None:
            _Lib._cdll = _load()
        return getattr(_Lib._cdll, name)


lib = _Lib()


def jit_policy(policy: int = -1) -> int:
    """vg_jit_policy: when prepared queries get a specialised single-pass
    kernel (0 never; 1, the default, queries that are not light, over
    tables of at least 8192 rows per compute unit; 2 all); returns the
    previous policy, and only reads it when policy < 0."""
    return lib.vg_jit_policy(policy)


def struct_sizes() -> list[int]:
    """virg_gpu_getsizes(): sizes of the 15 ABI structs as compiled."""
    p = lib.virg_gpu_getsizes()
    return [p[i] for i in range(15)]


class Virginian:
    """One database instance (virg_init .. virg_close)."""

    def __init__(self) -> None:
        self._st = State()
        if lib.virg_init(C.byref(self._st)) != VIRG_SUCCESS:
            raise VirginianError("virg_init failed")
        self.ptr = C.pointer(self._st)
        self.use_gpu = True

    def close(self) -> None:
        if self.ptr is not None:
            ok = lib.virg_close(self.ptr)
            self.ptr = None
            if ok != VIRG_SUCCESS:
                raise VirginianError("virg_close failed (the database file was not written)")

    # ---- database files (the reference's format) ----------------------------
    def create_db(self, path: str) -> None:
        """Back this (still empty) database with a new file (virg_db_create)."""
        if lib.virg_db_create(self.ptr, str(path).encode()) != VIRG_SUCCESS:
            raise VirginianError(f"virg_db_create failed for {path}")

    def open_db(self, path: str) -> None:
        """Open a database file, before any table is in memory (virg_db_open)."""
        if lib.virg_db_open(self.ptr, str(path).encode()) != VIRG_SUCCESS:
            raise VirginianError(f"virg_db_open failed for {path}")

    def tablet_limit(self, tablets: int) -> int:
        """Tablets kept in host memory while a database file is open
        (vg_tablet_limit; 0 = no limit); returns the previous limit."""
        return lib.vg_tablet_limit(self.ptr, tablets)

    def pager_stats(self, tablet_id: int = 0) -> dict:
        """Resident tablets, evictions and reloads (vg_pager_stats)."""
        r, e, f, ri = C.c_uint(), C.c_ulonglong(), C.c_ulonglong(), C.c_int()
        lib.vg_pager_stats(self.ptr, C.byref(r), C.byref(e), C.byref(f), tablet_id, C.byref(ri))
        return {"resident": r.value, "evictions": e.value, "file_loads": f.value, "tablet_resident": bool(ri.value)}

    def close_db(self) -> None:
        """Write every tablet in memory and the meta information to the file
        and close it (virg_db_close); the instance then holds an empty
        database in memory."""
        if lib.virg_db_close(self.ptr) != VIRG_SUCCESS:
            raise VirginianError("virg_db_close failed")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def use_gpu(self) -> bool:
        return bool(self._st.use_gpu)

    @use_gpu.setter
    def use_gpu(self, on: bool) -> None:
        self._st.use_gpu = 1 if on else 0

    @property
    def use_stream(self) -> bool:
        """virginian.use_stream: virg_vm_gpu streams tablets from host
        memory instead of keeping them resident in HBM."""
        return bool(self._st.use_stream)

    @use_stream.setter
    def use_stream(self, on: bool) -> None:
        self._st.use_stream = 1 if on else 0

    @property
    def use_mmap(self) -> bool:
        """virginian.use_mmap: the reference's mapped mode -- the data stays
        in host memory (here it streams through HBM as with use_stream)."""
        return bool(self._st.use_mmap)

    @use_mmap.setter
    def use_mmap(self, on: bool) -> None:
        self._st.use_mmap = 1 if on else 0

    @property
    def state(self) -> State:
        return self._st

    def device(self) -> tuple[str, int]:
        name = C.create_string_buffer(64)
        cus = C.c_int(0)
        lib.vg_device_info(self.ptr, name, 64, C.byref(cus))
        return name.value.decode(), cus.value

    # ---- tables -----------------------------------------------------------
    def table_id(self, name: str) -> int:
        tid = C.c_uint(0)
        if lib.virg_table_getid(self.ptr, name.encode(), C.byref(tid)) != VIRG_SUCCESS:
            raise VirginianError(f"no table {name!r}")
        return tid.value

    def create_table(self, name: str, key_type: int, columns: list[tuple[str, int]]) -> int:
        if lib.virg_table_create(self.ptr, name.encode(), key_type) != VIRG_SUCCESS:
            raise VirginianError("virg_table_create failed")
        tid = self.table_id(name)
        for cname, ctype in columns:
            if lib.virg_table_addcolumn(self.ptr, tid, cname.encode(), ctype) != VIRG_SUCCESS:
                raise VirginianError(f"virg_table_addcolumn {cname} failed")
        return tid

    def append_tablet(self, table_id: int, key: np.ndarray, cols: list[np.ndarray],
                      capacity: int | None = None) -> None:
        n = len(key)
        cap = capacity if capacity is not None else max(16, (n + 15) // 16 * 16)
        key = np.ascontiguousarray(key)
        cols = [np.ascontiguousarray(c) for c in cols]
        meta = self.tablets(table_id)[0].contents
        if len(cols) != meta.fixed_columns or key.dtype.itemsize != meta.key_stride or any(
                len(c) != n or c.dtype.itemsize != meta.fixed_stride[j] for j, c in enumerate(cols)):
            raise VirginianError("append_tablet: arrays do not match the table's key and columns")
        arr = (C.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
        if lib.virg_table_appendtablet(self.ptr, table_id, cap, n, key.ctypes.data, arr) != VIRG_SUCCESS:
            raise VirginianError("virg_table_appendtablet failed")

    def insert(self, table_id: int, key: np.ndarray, cols: list[np.ndarray]) -> None:
        """Row-at-a-time insert (reference path: 8 MB tablets)."""
        rowdt = np.dtype([(f"c{i}", c.dtype) for i, c in enumerate(cols)])
        rows = np.empty(len(key), dtype=rowdt)
        for i, c in enumerate(cols):
            rows[f"c{i}"] = c
        key = np.ascontiguousarray(key)
        ks = key.dtype.itemsize
        for r in range(len(key)):
            if lib.virg_table_insert(self.ptr, table_id, key.ctypes.data + r * ks,
                                     rows.ctypes.data + r * rowdt.itemsize, None) != VIRG_SUCCESS:
                raise VirginianError("virg_table_insert failed")

    def gen_table(self, name: str, rows: int, tablet_rows: int = 1 << 20, seed: int = 1,
                  threads: int = 8, device: bool = False) -> int:
        """Standard 7-column synthetic table (see csrc/host/gen.c); with
        device=True built by the GPU (vg_gen_table), the same bits."""
        if device:
            if lib.vg_gen_table(self.ptr, name.encode(), rows, tablet_rows, seed) != VIRG_SUCCESS:
                raise VirginianError("vg_gen_table failed")
        elif lib.virg_gen_table(self.ptr, name.encode(), rows, tablet_rows, seed, threads) != VIRG_SUCCESS:
            raise VirginianError("virg_gen_table failed")
        return self.table_id(name)

    def tablets(self, table_id: int) -> list:
        n = lib.virg_table_tablets(self.ptr, table_id, None, 0)
        arr = (C.POINTER(TabletMeta) * max(n, 1))()
        lib.virg_table_tablets(self.ptr, table_id, arr, n)
        return [arr[i] for i in range(n)]

    def numrows(self, table_id: int) -> int:
        n = C.c_uint(0)
        lib.virg_table_numrows(self.ptr, table_id, C.byref(n))
        return n.value

    # ---- programs ---------------------------------------------------------
    def compile(self, sql: str):
        """SQL -> virg_vm (caller frees with free_vm)."""
        vm = lib.virg_vm_init()
        err = C.create_string_buffer(256)
        if lib.virg_sql_compile(self.ptr, sql.encode(), vm, err, 256) != VIRG_SUCCESS:
            lib.virg_vm_cleanup(self.ptr, vm)
            raise VirginianError(err.value.decode())
        return vm

    def free_vm(self, vm) -> None:
        lib.virg_vm_cleanup(self.ptr, vm)

    def execute(self, vm) -> dict:
        """virg_vm_execute on the GPU, then the logical result as arrays."""
        if lib.virg_vm_execute(self.ptr, vm) != VIRG_SUCCESS:
            raise VirginianError("virg_vm_execute failed")
        return self.results(vm)

    def results(self, vm) -> dict:
        n = C.c_ulonglong(0)
        lib.virg_result_columns(self.ptr, vm, None, C.byref(n))
        names, types = result_schema(vm)
        outs = [np.empty(n.value, dtype=NP_TYPE[t]) for t in types]
        ptrs = (C.c_void_p * max(len(outs), 1))(*[o.ctypes.data for o in outs])
        lib.virg_result_columns(self.ptr, vm, ptrs, C.byref(n))
        return dict(zip(names, outs))

    def query(self, sql: str) -> dict:
        vm = self.compile(sql)
        try:
            return self.execute(vm)
        finally:
            self.free_vm(vm)

    def stream_select(self, sql_or_ops, capacity: int | None = None, batch_rows: int = 0,
                      pinned: bool = True) -> tuple[dict, StreamStats]:
        """Out-of-core select (vg_stream_select): the table stays in host
        memory and only referenced columns stream through HBM.  Results go
        to pinned host buffers unless pinned=False."""
        if isinstance(sql_or_ops, str):
            vm = self.compile(sql_or_ops)
            try:
                arr, n = ops_array(vm)
                names, types = result_schema(vm)
                arr = (Op * n)(*arr)
            finally:
                self.free_vm(vm)
            keep = None
        else:
            arr, n, keep = make_ops(sql_or_ops)
            names = [o[4] for o in sql_or_ops if o[0] == OP_ResultColumn]
            types = [o[1] for o in sql_or_ops if o[0] == OP_ResultColumn]
        tid = None
        for i in range(n):
            if arr[i].op == OP_Table:
                tid = arr[i].p1
                break
        cap = capacity if capacity is not None else self.numrows(tid)
        bufs, outs = [], []
        try:
            for t in types:
                dt = np.dtype(NP_TYPE[t])
                nbytes = max(cap * dt.itemsize, 1)
                if pinned:
                    ptr = lib.vg_host_alloc(nbytes)
                    if not ptr:
                        raise VirginianError("vg_host_alloc failed")
                    bufs.append(ptr)
                    outs.append(np.ctypeslib.as_array((C.c_uint8 * nbytes).from_address(ptr)).view(dt))
                else:
                    outs.append(np.empty(max(cap, 1), dtype=dt))
            ptrs = (C.c_void_p * max(len(outs), 1))(*[o.ctypes.data for o in outs])
            st = StreamStats()
            err = C.create_string_buffer(256)
            if lib.vg_stream_select(self.ptr, arr, n, ptrs, cap, batch_rows, C.byref(st), err, 256) != VIRG_SUCCESS:
                raise VirginianError(err.value.decode() or "vg_stream_select failed")
            cols = {nm: o[:st.rows_out].copy() for nm, o in zip(names, outs)}
        finally:
            for ptr in bufs:
                lib.vg_host_free(ptr)
        return cols, st

    def host_memory(self) -> tuple[int, int]:
        """(pinned, pageable) bytes of tablet slot memory allocated so far."""
        a, b = C.c_ulonglong(0), C.c_ulonglong(0)
        if lib.vg_host_memory(self.ptr, C.byref(a), C.byref(b)) != VIRG_SUCCESS:
            raise VirginianError("vg_host_memory failed")
        return a.value, b.value

    def pin_limit(self, nbytes: int) -> None:
        """Pin at most nbytes of tablet memory from now on (0: no limit)."""
        if lib.vg_host_pin_limit(self.ptr, nbytes) != VIRG_SUCCESS:
            raise VirginianError("vg_host_pin_limit failed")

    def link_probe(self, nbytes: int = 1 << 30, reps: int = 3) -> tuple[float, float]:
        """(host-to-device, device-to-host) GB/s of pinned copies."""
        h, d = C.c_double(0), C.c_double(0)
        if lib.vg_link_probe(self.ptr, nbytes, reps, C.byref(h), C.byref(d)) != VIRG_SUCCESS:
            raise VirginianError("vg_link_probe failed")
        return h.value, d.value

    def prepare(self, sql_or_ops, capacity: int = 0, flags: int = 0) -> "DeviceQuery":
        """flags: VG_QUERY_LIGHT / VG_QUERY_NO_OUTPUT (include/virg_gpu.h)."""
        return DeviceQuery(self, sql_or_ops, capacity, flags)

    def device_alloc(self, nbytes: int) -> int:
        ptr = lib.vg_device_alloc(self.ptr, nbytes)
        if not ptr:
            raise VirginianError(f"cannot allocate {nbytes} bytes of device memory")
        return ptr

    def device_free(self, ptr: int) -> None:
        lib.vg_device_free(self.ptr, ptr)


def program(vm) -> list[tuple]:
    """The op-program of a virg_vm as (op, p1, p2, p3, p4) tuples (p4 as
    int / float / str depending on the op)."""
    out = []
    for i in range(vm.contents.num_ops):
        o = vm.contents.stmt[i]
        if o.op == OP_ResultColumn:
            p4 = o.p4.s.decode() if o.p4.s else None
        elif o.op == OP_Float:
            p4 = o.p4.f
        else:
            p4 = o.p4.i
        out.append((o.op, o.p1, o.p2, o.p3, p4))
    return out


def result_schema(vm) -> tuple[list[str], list[int]]:
    names, types = [], []
    for i in range(vm.contents.num_ops):
        o = vm.contents.stmt[i]
        if o.op == OP_ResultColumn:
            names.append(o.p4.s.decode() if o.p4.s else f"c{len(names)}")
            types.append(o.p1)
    return names, types


def ops_array(vm) -> tuple:
    n = vm.contents.num_ops
    arr = (Op * n)()
    for i in range(n):
        arr[i] = vm.contents.stmt[i]
    return arr, n


def make_ops(ops: list[tuple]) -> tuple:
    """(op, p1, p2, p3, p4) tuples -> Op array; p4 may be int, float or str."""
    arr = (Op * len(ops))()
    keep = []
    for i, (op, p1, p2, p3, p4) in enumerate(ops):
        arr[i].op, arr[i].p1, arr[i].p2, arr[i].p3 = op, p1, p2, p3
        arr[i].p4.li = 0
        if isinstance(p4, float):
            arr[i].p4.f = p4
        elif isinstance(p4, str):
            b = C.create_string_buffer(p4.encode(), VIRG_MAX_COLUMN_NAME)
            keep.append(b)
            arr[i].p4.s = C.cast(b, C.c_char_p)
        elif p4 is not None:
            arr[i].p4.i = int(p4)
    return arr, len(ops), keep


class DeviceQuery:
    """A query prepared for repeated execution with its table resident in
    HBM (vg_query_* in include/virg_gpu.h)."""

    def __init__(self, db: Virginian, sql_or_ops, capacity: int = 0, flags: int = 0):
        self.db = db
        if isinstance(sql_or_ops, str):
            vm = db.compile(sql_or_ops)
            try:
                arr, n = ops_array(vm)
                self.names, self.types = result_schema(vm)
                self._keep = []
            finally:
                db.free_vm(vm)
        else:
            arr, n, self._keep = make_ops(sql_or_ops)
            self.names = [o[4] if isinstance(o[4], str) else f"c{k}"
                          for k, o in enumerate([o for o in sql_or_ops if o[0] == OP_ResultColumn])]
            self.types = [o[1] for o in sql_or_ops if o[0] == OP_ResultColumn]
        self._ops = arr
        err = C.create_string_buffer(256)
        self.h = lib.vg_query_create(db.ptr, arr, n, capacity, flags, err, 256)
        if not self.h:
            raise VirginianError(err.value.decode() or "vg_query_create failed")
        rows, tiles, rb, wb, nout = (C.c_ulonglong(), C.c_uint(), C.c_uint(), C.c_uint(), C.c_uint())
        lib.vg_query_info(self.h, C.byref(rows), C.byref(tiles), C.byref(rb), C.byref(wb), C.byref(nout))
        self.rows, self.tiles = rows.value, tiles.value
        self.read_bytes_per_row, self.write_bytes_per_row = rb.value, wb.value
        self.last_rows = 0

    def run(self) -> int:
        n = C.c_ulonglong(0)
        if lib.vg_query_run(self.h, C.byref(n)) != VIRG_SUCCESS:
            raise VirginianError("vg_query_run failed")
        self.last_rows = n.value
        return n.value

    def out_widths(self) -> list[int]:
        w = (C.c_uint * VIRG_MAX_COLUMNS)()
        n = lib.vg_query_out_widths(self.h, w)
        return list(w)[:n]

    def set_outputs(self, ptrs: list[int], capacity: int) -> None:
        arr = (C.c_void_p * max(len(ptrs), 1))(*ptrs)
        lib.vg_query_set_outputs(self.h, arr, capacity)
        self.capacity = capacity

    def sync(self) -> int:
        """Wait for the last launch of this query; returns its rows."""
        n = C.c_ulonglong(0)
        if lib.vg_query_sync(self.h, C.byref(n)) != VIRG_SUCCESS:
            raise VirginianError("vg_query_sync failed")
        self.last_rows = n.value
        return n.value

    def fetch(self) -> dict:
        outs = [np.empty(self.last_rows, dtype=NP_TYPE[t]) for t in self.types]
        ptrs = (C.c_void_p * max(len(outs), 1))(*[o.ctypes.data for o in outs])
        if lib.vg_query_fetch(self.h, ptrs, self.last_rows) != VIRG_SUCCESS:
            raise VirginianError("vg_query_fetch failed")
        return dict(zip(self.names, outs))

    def time(self, warmup: int, iters: int) -> tuple[float, int]:
        ms = C.c_double(0)
        n = C.c_ulonglong(0)
        if lib.vg_query_time(self.h, warmup, iters, C.byref(ms), C.byref(n)) != VIRG_SUCCESS:
            raise VirginianError("vg_query_time failed")
        self.last_rows = n.value
        return ms.value, n.value

    def kernel(self) -> dict:
        """What runs this query (vg_query_kernel): the specialised single-pass
        kernel (fused) or the interpreter's two passes."""
        f, r, sc, g, ms, m = C.c_int(), C.c_int(), C.c_uint(), C.c_uint(), C.c_double(), C.c_int()
        lib.vg_query_kernel(self.h, C.byref(f), C.byref(r), C.byref(sc), C.byref(g), C.byref(ms), C.byref(m))
        return {"fused": bool(f.value), "rows_per_lane": r.value, "tiles_per_chunk": sc.value,
                "grid": g.value, "compile_ms": ms.value, "aligned": bool(m.value & 1), "staged": bool(m.value & 2),
                "held": bool(m.value & 4), "pipelined": bool(m.value & 8), "waves": (m.value >> 8) & 0xFF}

    def time_passes(self, warmup: int, iters: int) -> tuple[float, float, int]:
        """(count-pass ms, write-pass ms, rows) from HIP events on the query's stream."""
        c, w = C.c_double(0), C.c_double(0)
        n = C.c_ulonglong(0)
        if lib.vg_query_time_passes(self.h, warmup, iters, C.byref(c), C.byref(w), C.byref(n)) != VIRG_SUCCESS:
            raise VirginianError("vg_query_time_passes failed")
        self.last_rows = n.value
        return c.value, w.value, n.value

    def close(self) -> None:
        if self.h:
            lib.vg_query_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GraphBatch:
    """Prepared queries replayed together as one HIP graph (vg_batch_* in
    include/virg_gpu.h); each query keeps its own results."""

    def __init__(self, queries: list[DeviceQuery]):
        self.queries = list(queries)
        arr = (C.c_void_p * max(len(self.queries), 1))(*[q.h for q in self.queries])
        err = C.create_string_buffer(256)
        self.h = lib.vg_batch_create(arr, len(self.queries), err, 256)
        if not self.h:
            raise VirginianError(err.value.decode() or "vg_batch_create failed")

    def run(self) -> list[int]:
        rows = (C.c_ulonglong * max(len(self.queries), 1))()
        if lib.vg_batch_launch(self.h) != VIRG_SUCCESS or lib.vg_batch_sync(self.h, rows) != VIRG_SUCCESS:
            raise VirginianError("batch run failed")
        out = list(rows)[:len(self.queries)]
        for q, n in zip(self.queries, out):
            q.last_rows = n
        return out

    def time(self, warmup: int, iters: int) -> float:
        ms = C.c_double(0)
        if lib.vg_batch_time(self.h, warmup, iters, C.byref(ms)) != VIRG_SUCCESS:
            raise VirginianError("vg_batch_time failed")
        return ms.value

    def close(self) -> None:
        if self.h:
            lib.vg_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _handles(queries) -> C.Array:
    return (C.c_void_p * max(len(queries), 1))(*[q.h for q in queries])


class PersistentBatch:
    """Prepared queries run by persistent launches over device item lists,
    captured as one HIP graph (vg_pbatch_* in include/virg_gpu.h): the
    interpreter's two passes, its single pass (single=1), or one launch of
    the batch's programs compiled together (jit=1)."""

    def __init__(self, queries: list[DeviceQuery], tiles_per_item: int = 0, order: int = 0,
                 grid_div: int = 0, split: int = 0, single: int = 0, direct: int = 0, db=None, jit: int = 0):
        self.queries = list(queries)
        cfg = PBatchConfig(tiles_per_item, grid_div, split, order, direct,
                           C.cast(db.ptr, C.c_void_p) if db else None, single, jit)
        err = C.create_string_buffer(256)
        self.h = lib.vg_pbatch_create(_handles(self.queries), len(self.queries), C.byref(cfg), err, 256)
        if not self.h:
            raise VirginianError(err.value.decode() or "vg_pbatch_create failed")

    def configure(self, tiles_per_item: int = 0, order: int = 0, grid_div: int = 0, split: int = 0,
                  single: int = 0, direct: int = 0, jit: int = 0) -> None:
        cfg = PBatchConfig(tiles_per_item, grid_div, split, order, direct, None, single, jit)
        err = C.create_string_buffer(256)
        if lib.vg_pbatch_configure(self.h, C.byref(cfg), err, 256) != VIRG_SUCCESS:
            raise VirginianError(err.value.decode() or "vg_pbatch_configure failed")

    def config(self) -> dict:
        """The configuration in use, with the number of query groups."""
        cfg = PBatchConfig()
        ng = C.c_uint(0)
        lib.vg_pbatch_get_config(self.h, C.byref(cfg), C.byref(ng))
        d = cfg.as_dict()
        d["groups"] = ng.value
        return d

    def run(self) -> tuple[list[int], list[float]]:
        """(rows per query, latency ms per query) of one run."""
        n = len(self.queries)
        rows = (C.c_ulonglong * max(n, 1))()
        lat = (C.c_double * max(n, 1))()
        if lib.vg_pbatch_launch(self.h) != VIRG_SUCCESS or lib.vg_pbatch_sync(self.h, rows, lat) != VIRG_SUCCESS:
            raise VirginianError("batch run failed")
        for q, r in zip(self.queries, rows):
            q.last_rows = r
        return list(rows)[:n], list(lat)[:n]

    def reload(self, queries: list[DeviceQuery]) -> None:
        """Run another set of queries from now on (after the last run)."""
        err = C.create_string_buffer(256)
        if lib.vg_pbatch_reload(self.h, _handles(queries), len(queries), err, 256) != VIRG_SUCCESS:
            raise VirginianError(err.value.decode() or "vg_pbatch_reload failed")
        self.queries = list(queries)

    def launch(self) -> None:
        if lib.vg_pbatch_launch(self.h) != VIRG_SUCCESS:
            raise VirginianError("batch launch failed")

    def done(self) -> bool:
        return lib.vg_pbatch_done(self.h) == 1

    def collect(self) -> tuple[list[int], list[float], list[int]]:
        """Wait for the launched run: (rows, latency ms, overflow flags)."""
        n = len(self.queries)
        rows = (C.c_ulonglong * max(n, 1))()
        lat = (C.c_double * max(n, 1))()
        flags = (C.c_uint * max(n, 1))()
        if lib.vg_pbatch_results(self.h, rows, lat, flags) != VIRG_SUCCESS:
            raise VirginianError("batch run failed")
        for q, r, f in zip(self.queries, rows, flags):
            q.last_rows = 0 if f else r
        return list(rows)[:n], list(lat)[:n], list(flags)[:n]

    def results(self) -> tuple[list[int], list[float], list[int]]:
        """Launch, wait, and return (rows, latency ms, overflow flags) per
        query without raising on a query whose result did not fit."""
        n = len(self.queries)
        rows = (C.c_ulonglong * max(n, 1))()
        lat = (C.c_double * max(n, 1))()
        flags = (C.c_uint * max(n, 1))()
        if lib.vg_pbatch_launch(self.h) != VIRG_SUCCESS or \
                lib.vg_pbatch_results(self.h, rows, lat, flags) != VIRG_SUCCESS:
            raise VirginianError("batch run failed")
        for q, r, f in zip(self.queries, rows, flags):
            q.last_rows = 0 if f else r
        return list(rows)[:n], list(lat)[:n], list(flags)[:n]

    def time(self, warmup: int, iters: int) -> float:
        ms = C.c_double(0)
        if lib.vg_pbatch_time(self.h, warmup, iters, C.byref(ms)) != VIRG_SUCCESS:
            raise VirginianError("vg_pbatch_time failed")
        return ms.value

    def tune(self, iters: int = 5) -> tuple[dict, float]:
        best = PBatchConfig()
        ms = C.c_double(0)
        err = C.create_string_buffer(256)
        if lib.vg_pbatch_tune(self.h, iters, C.byref(best), C.byref(ms), err, 256) != VIRG_SUCCESS:
            raise VirginianError(err.value.decode() or "vg_pbatch_tune failed")
        return best.as_dict(), ms.value

    def close(self) -> None:
        if self.h:
            lib.vg_pbatch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plain_time(queries: list[DeviceQuery], warmup: int, iters: int) -> tuple[float, list[float]]:
    """(ms per round, per-query latency ms of the last round) launching the
    queries one after another on one stream."""
    n = len(queries)
    ms = C.c_double(0)
    lat = (C.c_double * max(n, 1))()
    if lib.vg_plain_time(_handles(queries), n, warmup, iters, C.byref(ms), lat) != VIRG_SUCCESS:
        raise VirginianError("vg_plain_time failed")
    return ms.value, list(lat)[:n]
