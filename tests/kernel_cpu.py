"""CPU runs of the specialised kernels (test infrastructure).

A program's specialised kernel is generated by the library exactly as for
the GPU (vg_jit_program: decode, shape, immediates, source), then its
per-chunk function is compiled for the host with the stand-ins of
tests/kernel_cpu_shim.h and run over the host tablets, one workgroup of
256 lanes (coroutines of one thread) taking the chunks in order (so every
look-back finds its predecessors published).  A pipelined shape runs in the
pipelined kernel's order: each chunk's filter phase into the other staging
buffer before the previous chunk's result phase.  The kernel entry itself,
the chunk counter and the epoch handling are GPU-only and not exercised
here.

    cols, rows, flags, shape = select(db, prog)

returns the result columns (as DeviceQuery.fetch), the rows selected, the
VG_OVF_* bits and the kernel's shape; tests compare them with the oracle.
Compiled programs are cached by source text under the temporary directory.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import re
import subprocess
import tempfile
from pathlib import Path

import numpy as np

import virginian_amd as va

HERE = Path(__file__).resolve().parent
CXX = "/opt/rocm/llvm/bin/clang++"
CACHE = Path(tempfile.gettempdir()) / "vg_kernel_cpu"
MAX_COLS, MAX_UOPS, MAX_OUT = 17, 48, 16


class Shape(C.Structure):
    _fields_ = [("R", C.c_int), ("S", C.c_uint32), ("TR", C.c_uint32), ("fbytes_row", C.c_uint32),
                ("ftile_bytes", C.c_uint32), ("stage_off", C.c_uint32), ("stage_wave_bytes", C.c_uint32),
                ("lds_bytes", C.c_uint32), ("aligned", C.c_int), ("staged", C.c_int), ("pipe", C.c_int), ("held", C.c_int),
                ("waves", C.c_int)]


class TabDesc(C.Structure):
    _fields_ = [("col", C.c_uint64 * MAX_COLS), ("rows", C.c_uint32), ("first_tile", C.c_uint32),
                ("aligned", C.c_uint32), ("pad", C.c_uint32)]


class Tile(C.Structure):
    _fields_ = [("tab", C.c_uint32), ("row0", C.c_uint32)]


class Ctrl(C.Structure):
    _fields_ = [("total", C.c_uint64), ("overflow", C.c_uint32), ("pad", C.c_uint32 * 5)]


class LaunchArgs(C.Structure):
    _fields_ = [("tabs", C.c_void_p), ("tiles", C.c_void_p), ("num_tiles", C.c_uint32),
                ("nchunks", C.c_uint32), ("state", C.c_void_p), ("status", C.c_void_p), ("ctrl", C.c_void_p),
                ("capacity", C.c_uint64), ("spin_limit", C.c_uint32), ("pad", C.c_uint32),
                ("out", C.c_void_p * MAX_OUT), ("imm", (C.c_uint64 * 2) * MAX_UOPS),
                ("oimm", C.c_uint64 * MAX_OUT)]


va.lib.vg_jit_program.restype = C.c_long
va.lib.vg_jit_program.argtypes = [C.POINTER(va.Op), C.c_uint, C.c_uint, C.POINTER(va.TabletMeta), C.c_int,
                                  C.POINTER(Shape), C.POINTER(LaunchArgs), C.POINTER(C.c_uint),
                                  C.POINTER(C.c_uint), C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]

RUNNER = r"""
#include <vector>
kc_dim threadIdx;
kc_dim gridDim = {1, 1, 1};
kc_block *kc_blk;

/* the lanes as coroutines on this thread */
static ucontext_t kc_sched;
static ucontext_t kc_ctx[256];
static int kc_state[256];            /* 0 runnable, 1 waiting, 2 done */
static kc_barrier *kc_waiting_on[256];
static unsigned kc_cur;

void kc_barrier::wait()
{
	const unsigned me = kc_cur;
	if (++count == n) {
		count = 0;
		gen++;
		for (unsigned i = 0; i < 256; i++)
			if (kc_state[i] == 1 && kc_waiting_on[i] == this)
				kc_state[i] = 0;
		return;
	}
	kc_state[me] = 1;
	kc_waiting_on[me] = this;
	swapcontext(&kc_ctx[me], &kc_sched);
}

static const vgx_launch_args *kc_A;
static u8 *kc_lds;
static u32 kc_ovf;
static u32 s_cnt[2][16][4];
static u64 s_pre;
static unsigned kc_stage_off;
static int kc_pipe;

static void kc_lane_main()
{
	const u32 nc = kc_A->nchunks;
	if (!kc_pipe) {
		for (u32 c = 0; c < nc; c++) {
			const u32 o = p0::chunk(*kc_A, c, 0ull, s_cnt[0], &s_pre, kc_lds, [] {});
			kc_ovf |= o;
			__syncthreads();
		}
	} else {
		/* as the pipelined kernel: chunk c+1's filter phase into the other
		 * buffer, then chunk c's result phase */
		const unsigned buf = kc_stage_off / 2;
		p0::held H[2];           /* a held shape's registers for the two chunks (this lane's) */
		u32 b = 0, n0 = 0;
		u64 vb0 = 0;
		if (nc) {
			p0::stage(*kc_A, 0, kc_lds, H[0]);
			n0 = p0::phase1(*kc_A, 0, 0ull, s_cnt[0], kc_lds, vb0, H[0]);
		}
		for (u32 c = 0; c < nc; c++) {
			__syncthreads();
			u32 n1 = 0;
			u64 vb1 = 0;
			u8 *const buf1 = kc_lds + (b ^ 1u) * buf;
			if (c + 1 < nc) {
				p0::stage(*kc_A, c + 1, buf1, H[b ^ 1u]);
				n1 = p0::phase1(*kc_A, c + 1, 0ull, s_cnt[b ^ 1u], buf1, vb1, H[b ^ 1u]);
			}
			kc_ovf |= p0::phase2(*kc_A, c, 0ull, s_cnt[b], &s_pre, kc_lds + b * buf, kc_lds + kc_stage_off,
				vb0, n0, H[b], [] {});
			n0 = n1;
			vb0 = vb1;
			b ^= 1u;
		}
	}
	kc_state[kc_cur] = 2;
}

/* every chunk in order, one workgroup of 256 lanes; VG_OVF_* bits into
 * the control block.  Returns 0, or -1 when lanes wait for ever. */
extern "C" int kc_run(const vgx_launch_args *A, unsigned lds_bytes, unsigned stage_off, int pipe)
{
	kc_stage_off = stage_off;
	kc_pipe = pipe;
	kc_block blk;
	blk.all.n = 256;
	for (int w = 0; w < 4; w++)
		blk.wave[w].n = 64;
	kc_blk = &blk;
	std::vector<u64> ldsv(lds_bytes / 8 + 2, 0);
	kc_lds = (u8 *)ldsv.data();
	kc_A = A;
	kc_ovf = 0;
	const size_t stack = 256 * 1024;
	std::vector<char> stacks(256 * stack);
	for (unsigned t = 0; t < 256; t++) {
		getcontext(&kc_ctx[t]);
		kc_ctx[t].uc_stack.ss_sp = stacks.data() + t * stack;
		kc_ctx[t].uc_stack.ss_size = stack;
		kc_ctx[t].uc_link = &kc_sched;
		makecontext(&kc_ctx[t], kc_lane_main, 0);
		kc_state[t] = 0;
	}
	for (;;) {
		unsigned ran = 0, done = 0;
		for (unsigned t = 0; t < 256; t++) {
			if (kc_state[t] == 2) {
				done++;
				continue;
			}
			if (kc_state[t] != 0)
				continue;
			kc_cur = t;
			threadIdx = kc_dim{t, 0, 0};
			swapcontext(&kc_sched, &kc_ctx[t]);
			ran++;
		}
		if (done == 256)
			break;
		if (!ran) {
			fprintf(stderr, "kernel_cpu: lanes wait at a barrier the others cannot reach\n");
			return -1;
		}
	}
	A->ctrl->overflow |= kc_ovf;
	return 0;
}
"""


def _host_source(src: str) -> str:
    """The generated source with the host stand-ins: the fixed-width typedefs
    of the runtime compiler dropped (the host's are used), address spaces
    erased, the kernel entry cut off, the runner appended."""
    src = re.sub(r"^typedef unsigned char uint8_t;.*\n", "", src, count=1, flags=re.M)
    src = re.sub(r"^typedef unsigned long long uint64_t;.*\n", "", src, count=1, flags=re.M)
    src = re.sub(r"__attribute__\(\(address_space\(\d\)\)\)", "", src)
    # 64-bit words as the host's <stdint.h> spells them (the runtime
    # compiler's typedefs make the two spellings one type)
    src = src.replace("typedef unsigned long long u64; typedef long long i64;", "typedef uint64_t u64; typedef int64_t i64;")
    cut = src.find('\nextern "C" __global__')
    assert cut > 0, "kernel entry not found"
    return '#include "kernel_cpu_shim.h"\n' + src[:cut] + RUNNER


def _compile(text: str) -> C.CDLL:
    CACHE.mkdir(parents=True, exist_ok=True)
    h = hashlib.sha256(text.encode()).hexdigest()[:24]
    so = CACHE / f"k_{h}.so"
    if not so.exists():
        # per process: parallel test workers may compile the same source at once
        cpp = CACHE / f"k_{h}.{os.getpid()}.cpp"
        cpp.write_text(text)
        tmp = CACHE / f"k_{h}.{os.getpid()}.so"
        r = subprocess.run([CXX, "-std=c++17", "-O1", "-ffp-contract=off", "-fPIC", "-shared",
                            "-w", f"-I{HERE}", str(cpp), "-o", str(tmp)], capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(f"host compile of {cpp} failed:\n{r.stderr[-4000:]}")
        os.replace(tmp, so)
    lib = C.CDLL(str(so))
    lib.kc_run.argtypes = [C.POINTER(LaunchArgs), C.c_uint, C.c_uint, C.c_int]
    lib.kc_run.restype = C.c_int
    return lib


def _program(db, prog):
    """(ops array, count, keep-alive, names, types, table id)"""
    if isinstance(prog, str):
        vm = db.compile(prog)
        try:
            arr, n = va.ops_array(vm)
            names, types = va.result_schema(vm)
        finally:
            db.free_vm(vm)
        keep = []
    else:
        arr, n, keep = va.make_ops(prog)
        names = [o[4] if isinstance(o[4], str) else f"c{k}"
                 for k, o in enumerate([o for o in prog if o[0] == va.OP_ResultColumn])]
        types = [o[1] for o in prog if o[0] == va.OP_ResultColumn]
    tid = [arr[i].p1 for i in range(n) if arr[i].op == va.OP_Table][0]
    return arr, n, keep, names, types, tid


def _desc(t) -> TabDesc:
    m = t.contents
    base = C.addressof(m)
    d = TabDesc()
    aligned = True
    for c in range(m.fixed_columns):
        a = base + m.fixed_block + m.fixed_offset[c]
        d.col[c] = a
        aligned = aligned and a % 16 == 0
    d.col[MAX_COLS - 1] = base + m.key_block
    aligned = aligned and d.col[MAX_COLS - 1] % 16 == 0
    d.rows = m.rows
    d.aligned = 1 if aligned else 0
    return d


def select(db, prog, capacity: int = 0):
    arr, n, keep, names, types, tid = _program(db, prog)
    pc = [i for i in range(n) if arr[i].op == va.OP_Parallel][0] + 1
    tabs = db.tablets(tid)
    descs = [_desc(t) for t in tabs]
    aligned = all(d.aligned for d in descs)
    sh, A = Shape(), LaunchArgs()
    nout, widths = C.c_uint(0), (C.c_uint * MAX_OUT)()
    err = C.create_string_buffer(1024)
    L = va.lib.vg_jit_program(arr, n, pc, tabs[0], 1 if aligned else 0, C.byref(sh), C.byref(A), C.byref(nout),
                              widths, None, 0, err, 1024)
    assert L > 0, err.value.decode()
    buf = C.create_string_buffer(L + 1)
    va.lib.vg_jit_program(arr, n, pc, tabs[0], 1 if aligned else 0, C.byref(sh), C.byref(A), C.byref(nout),
                          widths, buf, L + 1, err, 1024)
    lib = _compile(_host_source(buf.value.decode()))
    tiles = [(i, r) for i, d in enumerate(descs) for r in range(0, d.rows, sh.TR)]
    rows = sum(d.rows for d in descs)
    cap = capacity or max(rows, 1)
    nchunks = (len(tiles) + sh.S - 1) // sh.S
    desc_arr = (TabDesc * max(len(descs), 1))(*descs)
    tile_arr = (Tile * max(len(tiles), 1))(*[Tile(a, b) for a, b in tiles])
    status = np.zeros(nchunks + 1, np.uint64)
    ctrl = Ctrl()
    outs = [np.zeros(cap * widths[j] + 16, np.uint8) for j in range(nout.value)]
    A.tabs = C.addressof(desc_arr)
    A.tiles = C.addressof(tile_arr)
    A.num_tiles = len(tiles)
    A.nchunks = nchunks
    A.status = status.ctypes.data
    A.ctrl = C.addressof(ctrl)
    A.capacity = cap
    A.spin_limit = 1 << 20
    for j, o in enumerate(outs):
        A.out[j] = o.ctypes.data
    assert lib.kc_run(C.byref(A), sh.lds_bytes, sh.stage_off, sh.pipe) == 0
    total = ctrl.total
    got = min(total, cap)
    cols = {}
    for j, (name, t) in enumerate(zip(names, types)):
        dt = np.dtype(va.NP_TYPE[t])
        assert dt.itemsize == widths[j], (name, dt, widths[j])
        cols[name] = outs[j][:got * widths[j]].view(dt).copy()
    shape = {f: getattr(sh, f) for f, _ in Shape._fields_}
    return cols, total, ctrl.overflow, shape
