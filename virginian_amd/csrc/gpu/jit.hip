/*
 * Prepare-time specialisation of a decoded program: the data-parallel
 * section of one query (the reference interprets it per row and per op,
 * src/vm/vm_gpu.cu:1152-1225 on the GPU, src/vm/virginia.c:228-891 on the
 * CPU) becomes a single-pass gfx950 kernel of its own, compiled with
 * hiprtc and cached by its source text.  Immediates are kernel arguments,
 * not literals, so queries differing only in constants (the selectivity
 * sweep) share one code object.
 *
 * The kernel (one launch per query run):
 *   - workgroups of 256 threads take chunks of S tiles of TR = 256*R rows
 *     from a per-query counter (one device-scope atomic per chunk: about
 *     15 ns each on MI355X, so chunks are ~16K rows);
 *   - phase 1 copies the filter columns of the chunk from HBM straight into
 *     LDS (global_load_lds), runs the filter micro-ops over them with each
 *     lane owning R rows (per-row masks, as the interpreter's rowstate) and
 *     counts the selected rows of every wave and tile;
 *   - the chunk's count is published at once and its output offset found
 *     by decoupled look-back over earlier chunks' status words;
 *   - phase 2, tile by tile: the other columns the result needs are loaded
 *     only for 4-row groups holding a selected row, the filter columns come
 *     from LDS again, the result micro-ops run, and each wave stages its
 *     selected values in LDS and writes them in row order with consecutive
 *     lanes on consecutive rows.
 * Every column is therefore read from HBM once.  Nothing is reset between
 * launches: the chunk counter is set back to 0 by the grab that completes
 * the launch's grabs (every workgroup makes exactly one grab past the last
 * chunk), which also moves the launch epoch on; status words carry the
 * epoch of the launch that wrote them, so the previous launch's words read
 * as unpublished.  Launches can therefore be captured in graphs and
 * replayed.  A workgroup only waits for chunks taken before its own, by
 * running workgroups, so the look-back cannot deadlock whatever else runs
 * on the device.
 */
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <stdint.h>
#include <string.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdarg.h>
#include <time.h>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

extern "C" {
#include "virginian.h"
#include "vg_prog.h"
#include "vgx.h"
}

/* ------------------------------------------------------------ prelude */
/* vg_prog.h and vgx_abi.h, as the build embedded them */
static const char *k_headers =
#include "vgx_headers.inc"
	;

static const char *k_prelude = R"VGX(
typedef unsigned char u8; typedef signed char i8; typedef unsigned short u16;
typedef unsigned int u32; typedef int i32; typedef unsigned long long u64; typedef long long i64;
typedef u32 u32x4 __attribute__((ext_vector_type(4)));
typedef u32 u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) void lds_void;
#define las __attribute__((address_space(3)))
typedef vg_tabdesc vgx_tab;
typedef vg_ctrl vgx_ctrl;
typedef vgx_launch_args vgx_args;
typedef const __attribute__((address_space(4))) vgx_args cvgx_args;
#define OOB 0x80000000u   /* a buffer offset past every resource: loads nothing, reads 0 */
#define ST_AGG (1ull << 62)
#define ST_PRE (2ull << 62)
#define ST_VBITS 38
#define ST_VAL ((1ull << ST_VBITS) - 1)
#define ST_EMASK ((1ull << 24) - 1)

__device__ __forceinline__ u32 uni(u32 x) { return __builtin_amdgcn_readfirstlane(x); }
/* a uniform read of read-only metadata through the constant address space
 * (a scalar load).  Scalars only: a struct copy is rewritten into loads
 * from the global address space, which become vector loads (a wait for
 * every load in flight before their values are used) */
template <typename T> __device__ __forceinline__ T ldc(const T *p)
{
	static_assert(__is_scalar(T), "ldc reads scalars; read a struct field by field");
	return *(const __attribute__((address_space(4))) T *)p;
}
template <typename T> __device__ __forceinline__ const __attribute__((address_space(1))) T *gp(const void *p)
{
	return (const __attribute__((address_space(1))) T *)p;
}
__device__ __forceinline__ u32 rank_below(u64 m)
{
	return __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
}
__device__ __forceinline__ u32 wsum(u32 x)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
	return x;
}
__device__ __forceinline__ u64 wsum64(u64 x)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
	return x;
}
/* every vector memory access of the wave done (loads into LDS included) */
__device__ __forceinline__ void wait_vm()
{
	__builtin_amdgcn_s_waitcnt(0x0F70);
	asm volatile("" ::: "memory");
}
/* LDS written by lanes of this wave visible to the others */
__device__ __forceinline__ void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
/* the interpreter's arithmetic: integers wrap, x / 0 = 0, INT_MIN / -1 wraps */
__device__ __forceinline__ i32 div_i32(i32 a, i32 b) { return b == 0 ? 0 : (b == -1 ? (i32)(0u - (u32)a) : a / b); }
__device__ __forceinline__ i64 div_i64(i64 a, i64 b) { return b == 0 ? 0 : (b == -1 ? (i64)(0ull - (u64)a) : a / b); }
__device__ __forceinline__ i32 div_chr(i32 a, i32 b) { return b == 0 ? 0 : a / b; }
__device__ __forceinline__ i64 mk64(u32 lo, u32 hi) { return (i64)(((u64)hi << 32) | lo); }
__device__ __forceinline__ double mkf64(u32 lo, u32 hi) { return __longlong_as_double(mk64(lo, hi)); }
__device__ __forceinline__ u32 lo64(i64 v) { return (u32)(u64)v; }
__device__ __forceinline__ u32 hi64(i64 v) { return (u32)((u64)v >> 32); }
__device__ __forceinline__ u32 lof64(double v) { return (u32)(u64)__double_as_longlong(v); }
__device__ __forceinline__ u32 hif64(double v) { return (u32)((u64)__double_as_longlong(v) >> 32); }
__device__ __forceinline__ u32 chr_bits(i32 v) { return (u32)(i32)(i8)v; }

/* Take the next chunk.  The launch makes nchunks + grid grabs in all; the
 * one that completes them sets the counter back to 0 and moves the epoch
 * on for the next launch (every workgroup read it before its first grab). */
__device__ __forceinline__ u32 grab(const vgx_args &A, u64 epoch)
{
	const u64 g = atomicAdd(&A.state[0], 1ull);
	if (g == (u64)A.nchunks + gridDim.x - 1) {
		__hip_atomic_store(&A.state[0], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		__hip_atomic_store(&A.state[1], (epoch + 1) & ST_EMASK, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}
	return g < A.nchunks ? (u32)g : 0xFFFFFFFFu;
}

/* Rows selected before chunk c (run by one whole wave, after the chunk's
 * own count was published with ST_AGG): the nearest running total of an
 * earlier chunk plus the single counts after it.  Publishes the running
 * total through chunk c.  *gave_up: a predecessor stayed unpublished for
 * spin_limit polls (the result is then unusable and flagged). */
__device__ u64 lookback(u64 *st, u32 c, u32 n, u32 lane, u64 epoch, u32 spin, bool *gave_up)
{
	const u64 etag = epoch << ST_VBITS;
	if (c == 0)
		return 0;
	u64 sum = 0;
	i64 j0 = (i64)c - 1;
	u32 polls = 0;
	for (;;) {
		const i64 j = j0 - (i64)lane;
		const u64 v = j >= 0 ? __hip_atomic_load(&st[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
		                     : (ST_PRE | etag);
		const bool cur = ((v >> ST_VBITS) & ST_EMASK) == epoch;
		const u32 f = cur ? (u32)(v >> 62) : 0u;
		const u64 pm = __ballot(f == 2u), nm = __ballot(f == 0u);
		const u64 first = pm & (~pm + 1);
		const u64 upto = pm ? (first | (first - 1)) : ~0ull;
		if (nm & upto) {
			if (++polls >= spin) {
				*gave_up = true;
				break;
			}
			/* back off: polls are device-scope loads that compete with
			 * the streaming loads of every other wave */
			if (polls < 4)
				__builtin_amdgcn_s_sleep(2);
			else
				__builtin_amdgcn_s_sleep(12);
			continue;
		}
		sum += wsum64(((upto >> lane) & 1ull) ? (v & ST_VAL) : 0ull);
		if (pm)
			break;
		j0 -= 64;
	}
	if (lane == 0)
		__hip_atomic_store(&st[c], ST_PRE | etag | ((sum + n) & ST_VAL), __ATOMIC_RELAXED,
			__HIP_MEMORY_SCOPE_AGENT);
	return sum;
}
)VGX";

/* ------------------------------------------------------------ codegen */
namespace {

struct out_buf {
	std::string s;
	void operator()(const char *fmt, ...) __attribute__((format(printf, 2, 3)))
	{
		char buf[1024];
		va_list ap;
		va_start(ap, fmt);
		const int n = vsnprintf(buf, sizeof buf, fmt, ap);
		va_end(ap);
		if (n < (int)sizeof buf) {
			s += buf;
			return;
		}
		std::string big((size_t)n + 1, '\0');
		va_start(ap, fmt);
		vsnprintf(&big[0], big.size(), fmt, ap);
		va_end(ap);
		big.resize((size_t)n);
		s += big;
	}
};

int tsize(int t)
{
	switch (t) {
	case VG_T_INT: case VG_T_FLOAT: return 4;
	case VG_T_INT64: case VG_T_DOUBLE: return 8;
	case VG_T_CHAR: return 1;
	}
	return 0;
}

/* C type of a value of virg_t t inside the kernel */
const char *ctype(int t)
{
	switch (t) {
	case VG_T_INT: return "i32";
	case VG_T_INT64: return "i64";
	case VG_T_FLOAT: return "float";
	case VG_T_DOUBLE: return "double";
	case VG_T_CHAR: return "i32";
	}
	return "void";
}

/* operand j (0: a, 1: b) of micro-op k for row r, as a value of type t */
std::string opnd(const vg_uop &u, int k, int j, int t, const char *r)
{
	const bool imm = j == 0 ? (u.flags & VG_F_AIMM) : (u.flags & VG_F_BIMM);
	const int lo = j == 0 ? u.a : u.b, hi = j == 0 ? u.a_hi : u.b_hi;
	char b[160];
	if (imm) {
		switch (t) {
		case VG_T_INT: case VG_T_CHAR: snprintf(b, sizeof b, "((i32)(u32)A.imm[%d][%d])", k, j); break;
		case VG_T_INT64: snprintf(b, sizeof b, "((i64)A.imm[%d][%d])", k, j); break;
		case VG_T_FLOAT: snprintf(b, sizeof b, "__uint_as_float((u32)A.imm[%d][%d])", k, j); break;
		case VG_T_DOUBLE: snprintf(b, sizeof b, "__longlong_as_double((i64)A.imm[%d][%d])", k, j); break;
		}
	} else {
		switch (t) {
		case VG_T_INT: case VG_T_CHAR: snprintf(b, sizeof b, "((i32)s%d[%s])", lo, r); break;
		case VG_T_INT64: snprintf(b, sizeof b, "mk64(s%d[%s], s%d[%s])", lo, r, hi, r); break;
		case VG_T_FLOAT: snprintf(b, sizeof b, "__uint_as_float(s%d[%s])", lo, r); break;
		case VG_T_DOUBLE: snprintf(b, sizeof b, "mkf64(s%d[%s], s%d[%s])", lo, r, hi, r); break;
		}
	}
	return b;
}

const char *zero_of(int t)
{
	switch (t) {
	case VG_T_FLOAT: return "0.0f";
	case VG_T_DOUBLE: return "0.0";
	}
	return "0";
}

/* store value expression x of type t into the slot(s) dst/dst_hi for row r,
 * only where cond holds (cond may be "true") */
void put(out_buf &o, int t, int dst, int dst_hi, const char *x, const char *cond, const char *ind)
{
	switch (t) {
	case VG_T_INT:
		o("%s{ const u32 v_ = (u32)(%s); s%d[r] = (%s) ? v_ : s%d[r]; }\n", ind, x, dst, cond, dst);
		break;
	case VG_T_FLOAT:
		o("%s{ const u32 v_ = __float_as_uint(%s); s%d[r] = (%s) ? v_ : s%d[r]; }\n", ind, x, dst, cond, dst);
		break;
	case VG_T_CHAR:
		o("%s{ const u32 v_ = chr_bits(%s); s%d[r] = (%s) ? v_ : s%d[r]; }\n", ind, x, dst, cond, dst);
		break;
	case VG_T_INT64:
		o("%s{ const i64 v_ = (%s); const bool c_ = (%s); s%d[r] = c_ ? lo64(v_) : s%d[r]; "
		  "s%d[r] = c_ ? hi64(v_) : s%d[r]; }\n", ind, x, cond, dst, dst, dst_hi, dst_hi);
		break;
	case VG_T_DOUBLE:
		o("%s{ const double v_ = (%s); const bool c_ = (%s); s%d[r] = c_ ? lof64(v_) : s%d[r]; "
		  "s%d[r] = c_ ? hif64(v_) : s%d[r]; }\n", ind, x, cond, dst, dst, dst_hi, dst_hi);
		break;
	}
}

/* the interpreter's convert<S, D> (a cast of register contents) */
std::string convert(int from, int to, const std::string &x)
{
	char b[256];
	if (to == VG_T_CHAR) {
		if (from == VG_T_FLOAT || from == VG_T_DOUBLE)
			snprintf(b, sizeof b, "((i32)(i8)(i32)(%s))", x.c_str());
		else
			snprintf(b, sizeof b, "((i32)(i8)(%s))", x.c_str());
	} else {
		snprintf(b, sizeof b, "((%s)(%s))", ctype(to), x.c_str());
	}
	return b;
}

/* Micro-ops [0, nuop) whose bit is set in `mask` (the interpreter's
 * run_uops), over the R rows of this lane; row masks valid, active, w0..w3. */
void gen_uops(out_buf &o, const vg_prog *P, uint32_t nuop, uint64_t mask)
{
	for (uint32_t k = 0; k < nuop; k++) {
		if (!((mask >> k) & 1ull))
			continue;
		const vg_uop &u = P->uop[k];
		o("\t\t\t/* micro-op %u (op %u) */\n", k, (unsigned)u.src_op);
		for (int j = 0; j < VG_NWAIT; j++)
			if ((u.join >> j) & 1u)
				o("\t\t\tactive |= w%d; w%d = 0;\n", j, j);
		const int t = u.type;
		if (u.kind == VG_K_INVALID) {
			o("\t\t\tvalid &= ~active; active = 0;\n");
			continue;
		}
		if (u.kind <= VG_K_NOT) {
			const std::string a = opnd(u, (int)k, 0, t, "r");
			const std::string b = u.kind == VG_K_NOT ? std::string() : opnd(u, (int)k, 1, t, "r");
			std::string e;
			const char *z = zero_of(t);
			switch (u.kind) {
			case VG_K_LT: e = a + " < " + b; break;
			case VG_K_LE: e = a + " <= " + b; break;
			case VG_K_GT: e = a + " > " + b; break;
			case VG_K_GE: e = a + " >= " + b; break;
			case VG_K_EQ: e = a + " == " + b; break;
			case VG_K_NE: e = a + " != " + b; break;
			case VG_K_AND: e = "(" + a + " != " + z + ") && (" + b + " != " + z + ")"; break;
			case VG_K_OR: e = "(" + a + " != " + z + ") || (" + b + " != " + z + ")"; break;
			case VG_K_NOT: e = "!(" + a + " != " + z + ")"; break;
			}
			o("\t\t\t{\n\t\t\t\tu32 c = 0;\n#pragma unroll\n\t\t\t\tfor (int r = 0; r < R; ++r)\n"
			  "\t\t\t\t\tc |= (%s) ? (1u << r) : 0u;\n", e.c_str());
			o("\t\t\t\tconst u32 taken = c & active;\n\t\t\t\tactive &= ~taken;\n");
			if (u.flags & VG_F_KEEP)
				o("\t\t\t\tw%d |= taken;\n", (int)u.wslot);
			else
				o("\t\t\t\tvalid &= ~taken;\n");
			o("\t\t\t}\n");
			continue;
		}
		const bool masked = (u.flags & VG_F_MASKED) != 0;
		const char *cond = masked ? "(active >> r) & 1u" : "true";
		std::string x;
		if (u.kind == VG_K_CAST) {
			x = convert(t, u.b, opnd(u, (int)k, 0, t, "r"));
			o("#pragma unroll\n\t\t\tfor (int r = 0; r < R; ++r)\n");
			put(o, u.b, u.dst, u.dst_hi, x.c_str(), "(active >> r) & 1u", "\t\t\t\t");
			continue;
		}
		const std::string a = opnd(u, (int)k, 0, t, "r");
		if (u.kind == VG_K_MOV) {
			x = a;
		} else {
			const std::string b = opnd(u, (int)k, 1, t, "r");
			const char *opc = u.kind == VG_K_ADD ? "+" : u.kind == VG_K_SUB ? "-" : u.kind == VG_K_MUL ? "*" : "/";
			switch (t) {
			case VG_T_INT:
				if (u.kind == VG_K_DIV)
					x = "div_i32(" + a + ", " + b + ")";
				else
					x = std::string("(i32)((u32)") + a + " " + opc + " (u32)" + b + ")";
				break;
			case VG_T_INT64:
				if (u.kind == VG_K_DIV)
					x = "div_i64(" + a + ", " + b + ")";
				else
					x = std::string("(i64)((u64)") + a + " " + opc + " (u64)" + b + ")";
				break;
			case VG_T_CHAR:
				if (u.kind == VG_K_DIV)
					x = "div_chr(" + a + ", " + b + ")";
				else
					x = std::string("(") + a + " " + opc + " " + b + ")";
				break;
			default:
				x = std::string("(") + a + " " + opc + " " + b + ")";
				break;
			}
		}
		o("#pragma unroll\n\t\t\tfor (int r = 0; r < R; ++r)\n");
		put(o, t, u.dst, u.dst_hi, x.c_str(), cond, "\t\t\t\t");
	}
}

/* where a result column's value comes from */
struct outsrc {
	int kind;        /* VG_OUT_* */
	int width;
	int lo, hi;      /* slot names (VG_OUT_SLOT) or emit-column variable index (VG_OUT_COLUMN) */
};

} /* namespace */

extern "C" int vgx_shape_of(const vg_prog *P, int aligned, int R_override, int S_override, int pipe, vgx_shape *sh,
	char *err, size_t errlen)
{
	memset(sh, 0, sizeof *sh);
	sh->aligned = aligned ? 1 : 0;
	/* pipe 2: a batch shape with the pipelined shape's staging budget and
	 * chunk size (half the staging), neither pipelined nor held */
	const bool small = pipe != 0, batch = pipe == 2;
	pipe = pipe == 1;
	sh->pipe = pipe ? 1 : 0;
	if (P->nuop > VG_MAX_UOPS || P->nout > VG_MAX_OUT) {
		snprintf(err, errlen, "program too large");
		return -1;
	}
	/* rows per lane from the register file size (values are held in VGPRs);
	 * a batch's chunk functions run at three workgroups per CU (168
	 * VGPRs), so they take fewer rows per lane from 8 slots on (1.316
	 * against 1.336 ms per run of the 256-program batch; from 6 slots:
	 * 1.396).  VG_JIT_BATCH_RSLOTS: the most slots at 16 rows per lane */
	const char *rse = getenv("VG_JIT_BATCH_RSLOTS");
	const uint32_t rs = !batch ? 10u : rse && *rse ? (uint32_t)atoi(rse) : 7u;
	int R = P->nslot <= rs ? 16 : P->nslot <= 2 * rs ? 8 : 4;
	if (R_override == 4 || R_override == 8 || R_override == 16)
		R = R_override;
	/* output staging: per wave, 256 rows x the widest group of outputs
	 * written together (groups of at most 16 bytes a row) */
	uint32_t stage_w = 0, gw = 0;
	for (uint32_t j = 0; j < P->nout; j++) {
		const uint32_t w = P->out[j].width;
		if (gw + w > 16u)
			gw = 0;
		gw += w;
		if (gw > stage_w)
			stage_w = gw;
	}
	const uint32_t stage_wave = 256u * (stage_w ? stage_w : 4u);
	/* filter columns staged in LDS: bytes per row; with the staging and
	 * the kernel's own LDS (under 1 KB) a workgroup takes at most 79 KB,
	 * so two fit a CU's 160 KB */
	uint32_t fbytes = 0, fwords = 0;
	for (uint32_t c = 0; c < P->ncol; c++)
		if (P->col[c].phase == VG_PHASE_FILTER) {
			fbytes += P->col[c].width == 1 ? 1u : P->col[c].width;
			fwords += P->col[c].width == 8 ? 2u : 1u;
		}
	const uint32_t two_per_cu = 79u * 1024u - 4u * stage_wave;
	uint32_t budget = two_per_cu < 64u * 1024u ? two_per_cu : 64u * 1024u;
	/* a pipelined kernel holds two chunks' filter columns at once */
	if (small)
		budget /= 2;
	const uint32_t min_rows = small ? 8192u : 16384u;
	/* a chunk is 16K rows (one grab and one look-back each); staging its
	 * filter columns in LDS saves reading them twice, when they fit at
	 * that size; otherwise they are loaded into registers tile by tile */
	const int R0 = R;
	while (R > 4 && fbytes * 256u * (uint32_t)R > budget)
		R /= 2;
	uint32_t TR = 256u * (uint32_t)R;
	uint32_t ftile = fbytes * TR;
	uint32_t S = 16384u / TR;
	if (ftile)
		while (S > 1 && ftile * S > budget)
			S--;
	const char *mode = getenv("VG_JIT_STAGE");
	int staged = S * TR >= min_rows && ftile <= budget;
	if (mode && (mode[0] == '0' || mode[0] == '1'))
		staged = mode[0] == '1' && ftile <= 96u * 1024u;
	if (!staged) {
		R = R0;
		TR = 256u * (uint32_t)R;
		ftile = 0;
		S = 16384u / TR;
	}
	/* a pipelined filter of one 4-byte (or 1-byte) column: held in
	 * registers instead, two chunks of 16K rows at a time (2 x 64 values a
	 * lane, S x R = 64 validity bits a lane), which reads the column once
	 * and needs no LDS for it.  Chunks of 16K rows at two workgroups per CU
	 * beat 8K rows at three (0.699 against 0.675 of HBM on the 64M-row
	 * sweep, the 0.1 % query 114 against 131 us): each chunk's chain of
	 * waits is paid half as often */
	int held = 0;
	const char *hm = getenv("VG_JIT_HELD");
	if (pipe && fwords == 1 && !(hm && hm[0] == '0') && !(mode && (mode[0] == '0' || mode[0] == '1'))) {
		held = 1;
		staged = 0;
		R = R0;
		TR = 256u * (uint32_t)R;
		ftile = 0;
		S = 16384u / TR;
	}
	if (S_override >= 1 && S_override <= 16)
		S = (uint32_t)S_override;
	if (S * (uint32_t)R > 64u) {
		snprintf(err, errlen, "chunk of %u tiles of %u rows per lane is too large", S, (unsigned)R);
		return -1;
	}
	sh->R = R;
	sh->S = S;
	sh->TR = TR;
	sh->fbytes_row = fbytes;
	sh->ftile_bytes = ftile;
	sh->staged = staged;
	sh->held = held;
	/* A launch bound of three waves per SIMD for a held shape
	 * (VG_JIT_WAVES=3): with 8K-row chunks (VG_JIT_S) it fits 168 VGPRs
	 * with 24 B spilled and runs three workgroups per CU (0.675, against
	 * 0.658 unbounded); with the default 16K-row chunks it would spill
	 * 288 B a lane and is built again unbounded.  No other bound is used:
	 * a register-tile kernel built for four waves per SIMD (128 VGPRs, 8 B
	 * spilled) reported a capacity overflow on the 64M-row benchmark
	 * query, where the same source at its own occupancy is exact. */
	const char *wv = getenv("VG_JIT_WAVES");
	sh->waves = held && wv && wv[0] == '3' ? 3 : 0;
	sh->stage_off = ftile * S * (pipe ? 2u : 1u);
	sh->stage_wave_bytes = stage_wave;
	sh->lds_bytes = sh->stage_off + 4u * sh->stage_wave_bytes;
	return 0;
}

/* Program P at shape sh as namespace `ns`: its constants and chunk(), the
 * work of one chunk (both phases, the look-back, the writes), returning
 * its VG_OVF_* bits.  chunk() calls grab_next() in thread 0 as the chunk's
 * last tile starts, for the caller's next chunk.  It is inlined into the
 * kernel that calls it, or, `outlined`, a function of its own (a module of
 * many programs then compiles in time linear in their number); its LDS
 * arguments then carry their address space. */
static void gen_program(out_buf &o, const vg_prog *P, const vgx_shape *sh, const char *ns, bool outlined = false)
{
	const int R = sh->R, NQ = R / 4;
	o("\nnamespace %s {\n", ns);
	o("constexpr int R = %d, NQ = %d;\nconstexpr u32 TR = %uu, S_ = %uu;\nconstexpr bool AL = %s;\n", R, NQ, sh->TR,
	  sh->S, sh->aligned ? "true" : "false");
	o("constexpr u32 RMASK = %s;\n", R == 32 ? "0xFFFFFFFFu" : R == 16 ? "0xFFFFu" : R == 8 ? "0xFFu" : "0xFu");

	/* column variables: loaded columns live in their slots; emit-only
	 * columns get variables of their own (e<c>, e<c>h) */
	std::vector<int> is_filter(P->ncol), need2(P->ncol);
	for (uint32_t c = 0; c < P->ncol; c++) {
		const vg_col &cl = P->col[c];
		is_filter[c] = cl.phase == VG_PHASE_FILTER;
		need2[c] = 0;
	}
	/* columns phase 2 needs: what the result micro-ops read (wload), every
	 * loaded column when the program is rerun whole, emit-only columns an
	 * output names */
	for (uint32_t c = 0; c < P->nload; c++)
		if (P->wfull || P->col[c].wload)
			need2[c] = 1;
	std::vector<outsrc> os(P->nout);
	for (uint32_t j = 0; j < P->nout; j++) {
		const vg_out &ou = P->out[j];
		os[j].kind = ou.kind;
		os[j].width = ou.width;
		if (ou.kind == VG_OUT_SLOT) {
			os[j].lo = ou.slot;
			os[j].hi = ou.slot_hi;
		} else if (ou.kind == VG_OUT_COLUMN) {
			os[j].lo = ou.col;
			need2[ou.col] = 1;
		}
	}
	/* LDS offsets of the filter columns inside a tile's staging area */
	std::vector<uint32_t> foff(P->ncol, 0);
	{
		uint32_t off = 0;
		for (uint32_t c = 0; c < P->ncol; c++)
			if (is_filter[c]) {
				foff[c] = off;
				off += sh->TR * (P->col[c].width == 1 ? 1u : P->col[c].width);
			}
	}

	/* --- helpers specific to this program --- */
	/* slot declarations (all zeroed: rows that never wrote a slot read 0) */
	out_buf decl;
	decl("\t\t\tu32");
	for (uint32_t s = 0; s < P->nslot; s++)
		decl("%s s%u[R]", s ? "," : "", s);
	if (!P->nslot)
		decl(" s_unused[1]");
	decl(";\n");
	for (uint32_t s = 0; s < P->nslot; s++)
		decl("#pragma unroll\n\t\t\tfor (int r = 0; r < R; ++r) s%u[r] = 0u;\n", s);
	for (uint32_t c = 0; c < P->ncol; c++)
		if (P->col[c].phase == VG_PHASE_EMIT && need2[c])
			decl("\t\t\tu32 e%u[R], e%uh[R];\n#pragma unroll\n\t\t\tfor (int r = 0; r < R; ++r) { e%u[r] = 0u; e%uh[r] = 0u; }\n",
				c, c, c, c);

	/* A chunk is worked in three steps, each a function: stage() issues
	 * the loads of the chunk's filter columns into the LDS staging buffer
	 * `lds` (staged shapes; nothing waits), phase1() evaluates the filter,
	 * counts and publishes the chunk's count, phase2() finds the chunk's
	 * output offset by look-back and writes its rows.  chunk() runs the
	 * three in a row; a pipelined kernel runs phase1() of its next chunk
	 * before phase2() of this one. */
	const char *pre_fn =
		"\tconst u32 tid = threadIdx.x, lane = tid & 63u, wave = uni(tid >> 6);\n"
		"\t(void)tid; (void)lane; (void)wave;\n"
		"\t{\n"
		"\t\tconst u32 t0 = c * S_;\n\t\tconst u32 nt = uni(min((u32)S_, A.num_tiles - t0));\n"
		"\t\t(void)t0; (void)nt;\n";
	/* a held shape's registers for one chunk's filter column(s) */
	o("struct held {");
	if (sh->held) {
		for (uint32_t c = 0; c < P->ncol; c++)
			if (is_filter[c]) {
				o(" u32 f%u[S_][R];", c);
				if (P->col[c].width == 8)
					o(" u32 f%uh[S_][R];", c);
			}
	} else {
		o(" u32 unused;");
	}
	o(" };\n");
	if (!sh->held) {
	o("template <typename AT>\n__device__ __attribute__((always_inline)) inline void stage(AT &A, const u32 c, u8 *lds, held &H)\n{\n");
	o("\t(void)H;\n");
	o("%s", pre_fn);
	}

	/* tile lookup */
	const char *tile_info =
		"\t\t\tconst vgx_tile ti = {ldc(&A.tiles[t0 + s].tab), ldc(&A.tiles[t0 + s].row0)};\n"
		"\t\t\tconst vgx_tab *tb = A.tabs + ti.tab;\n"
		"\t\t\tconst u32 trows = ldc(&tb->rows);\n"
		"\t\t\tconst u32 nrows = uni(min((u32)TR, trows - ti.row0));\n"
		"\t\t\tconst bool fast = nrows == TR && ldc(&tb->aligned) != 0;\n";

	/* phase 1: stage filter columns into LDS */
	if (sh->staged) {
	o("\t\t/* phase 1: the chunk's filter columns, HBM -> LDS */\n");
	o("\t\tfor (u32 s = 0; s < nt; ++s) {\n%s", tile_info);
	o("\t\t\tu8 *const st = lds + s * %uu;\n", sh->ftile_bytes);
	for (uint32_t c = 0; c < P->ncol; c++) {
		if (!is_filter[c])
			continue;
		const vg_col &cl = P->col[c];
		o("\t\t\t{\n\t\t\t\tconst char *base = (const char *)ldc(&tb->col[%u]) + (u64)ti.row0 * %u;\n", cl.src, cl.width);
		o("\t\t\t\tu8 *const d = st + %u;\n", foff[c]);
		o("#pragma unroll\n\t\t\t\tfor (int q = 0; q < NQ; ++q) {\n"
		  "\t\t\t\t\tconst u32 r0 = wave * (NQ * 256u) + q * 256u;\n");
		if (cl.width == 4) {
			o("\t\t\t\t\tif (fast) {\n\t\t\t\t\t\t__builtin_amdgcn_global_load_lds((const void *)(base + (u64)(r0 + lane * 4u) * 4u), (lds_void *)(d + r0 * 4u), 16, 0, 0);\n"
			  "\t\t\t\t\t} else {\n\t\t\t\t\t\tu32x4 v = {0u, 0u, 0u, 0u};\n\t\t\t\t\t\tconst u32 rr = r0 + lane * 4u;\n"
			  "\t\t\t\t\t\tfor (int j = 0; j < 4; ++j) if (rr + j < nrows) v[j] = gp<u32>(base)[rr + j];\n"
			  "\t\t\t\t\t\t*(u32x4 *)(d + rr * 4u) = v;\n\t\t\t\t\t}\n");
		} else if (cl.width == 8) {
			o("\t\t\t\t\tif (fast) {\n"
			  "\t\t\t\t\t\t__builtin_amdgcn_global_load_lds((const void *)(base + (u64)r0 * 8u + lane * 16u), (lds_void *)(d + r0 * 8u), 16, 0, 0);\n"
			  "\t\t\t\t\t\t__builtin_amdgcn_global_load_lds((const void *)(base + (u64)r0 * 8u + 1024u + lane * 16u), (lds_void *)(d + r0 * 8u + 1024u), 16, 0, 0);\n"
			  "\t\t\t\t\t} else {\n\t\t\t\t\t\tconst u32 rr = r0 + lane * 4u;\n"
			  "\t\t\t\t\t\tfor (int j = 0; j < 4; ++j) {\n\t\t\t\t\t\t\tu32x2 v = {0u, 0u};\n"
			  "\t\t\t\t\t\t\tif (rr + j < nrows) v = gp<u32x2>(base)[rr + j];\n"
			  "\t\t\t\t\t\t\t*(u32x2 *)(d + (rr + j) * 8u) = v;\n\t\t\t\t\t\t}\n\t\t\t\t\t}\n");
		} else {
			o("\t\t\t\t\tif (fast) {\n\t\t\t\t\t\t__builtin_amdgcn_global_load_lds((const void *)(base + r0 + lane * 4u), (lds_void *)(d + r0), 4, 0, 0);\n"
			  "\t\t\t\t\t} else {\n\t\t\t\t\t\tu32 v = 0u;\n\t\t\t\t\t\tconst u32 rr = r0 + lane * 4u;\n"
			  "\t\t\t\t\t\tfor (int j = 0; j < 4; ++j) if (rr + j < nrows) v |= (u32)gp<u8>(base)[rr + j] << (8 * j);\n"
			  "\t\t\t\t\t\t*(u32 *)(d + rr) = v;\n\t\t\t\t\t}\n");
		}
		o("\t\t\t\t}\n\t\t\t}\n");
	}
	o("\t\t}\n");
	}
	if (!sh->held)
		o("\t}\n}\n\n");

	/* read filter column c of tile s from LDS into its slots (this lane's rows) */
	auto lds_read = [&](out_buf &b, uint32_t c, const char *ind) {
		const vg_col &cl = P->col[c];
		b("%s{\n%s\tconst u8 *const d = lds + s * %uu + %u;\n", ind, ind, sh->ftile_bytes, foff[c]);
		b("#pragma unroll\n%s\tfor (int q = 0; q < NQ; ++q) {\n%s\t\tconst u32 rr = wave * (NQ * 256u) + q * 256u + lane * 4u;\n", ind, ind);
		if (cl.width == 4) {
			b("%s\t\tconst u32x4 v = *(const u32x4 *)(d + rr * 4u);\n", ind);
			b("%s\t\ts%u[4 * q] = v.x; s%u[4 * q + 1] = v.y; s%u[4 * q + 2] = v.z; s%u[4 * q + 3] = v.w;\n",
				ind, cl.slot, cl.slot, cl.slot, cl.slot);
		} else if (cl.width == 8) {
			b("%s\t\tconst u32x4 v0 = *(const u32x4 *)(d + rr * 8u), v1 = *(const u32x4 *)(d + rr * 8u + 16u);\n", ind);
			b("%s\t\ts%u[4 * q] = v0.x; s%u[4 * q] = v0.y; s%u[4 * q + 1] = v0.z; s%u[4 * q + 1] = v0.w;\n",
				ind, cl.slot, cl.slot_hi, cl.slot, cl.slot_hi);
			b("%s\t\ts%u[4 * q + 2] = v1.x; s%u[4 * q + 2] = v1.y; s%u[4 * q + 3] = v1.z; s%u[4 * q + 3] = v1.w;\n",
				ind, cl.slot, cl.slot_hi, cl.slot, cl.slot_hi);
		} else {
			b("%s\t\tconst u32 v = *(const u32 *)(d + rr);\n", ind);
			b("%s\t\tfor (int j = 0; j < 4; ++j) s%u[4 * q + j] = (u32)(i32)(i8)(v >> (8 * j));\n", ind, cl.slot);
		}
		b("%s\t}\n%s}\n", ind, ind);
	};

	/* HBM loads of column c for the 4-row groups with a row in `sel`: buffer loads whose offset is out of range for the other
	 * groups (they read nothing and return 0), so no lane branches and all
	 * loads in flight together; tiles past a tablet's last row are cut off
	 * by the resource's size */
	auto col_load = [&](uint32_t c, const char *tlo = nullptr, const char *thi = nullptr) {
		const vg_col &cl = P->col[c];
		const bool emit = cl.phase == VG_PHASE_EMIT;
		char lo[64], hi[64];
		if (tlo) {
			snprintf(lo, sizeof lo, "%s", tlo);
			snprintf(hi, sizeof hi, "%s", thi ? thi : tlo);
		} else if (emit) {
			snprintf(lo, sizeof lo, "e%u", c);
			snprintf(hi, sizeof hi, "e%uh", c);
		} else {
			snprintf(lo, sizeof lo, "s%u", cl.slot);
			snprintf(hi, sizeof hi, "s%u", cl.slot_hi);
		}
		o("\t\t\t{\n\t\t\t\tconst __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc("
		  "(void *)((const char *)ldc(&tb->col[%u]) + (u64)ti.row0 * %u), 0, (int)(nrows * %uu), 0x00020000);\n",
		  cl.src, cl.width, cl.width);
		o("#pragma unroll\n\t\t\t\tfor (int q = 0; q < NQ; ++q) {\n"
		  "\t\t\t\t\tconst u32 g = (sel >> (4 * q)) & 0xFu;\n"
		  "\t\t\t\t\tconst u32 rr = wave * (NQ * 256u) + q * 256u + lane * 4u;\n");
		if (cl.width == 4) {
			o("\t\t\t\t\tif constexpr (AL) {\n\t\t\t\t\tconst u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, g ? rr * 4u : OOB, 0, 0);\n"
			  "\t\t\t\t\t%s[4 * q] = v.x; %s[4 * q + 1] = v.y; %s[4 * q + 2] = v.z; %s[4 * q + 3] = v.w;\n"
			  "\t\t\t\t\t} else {\n#pragma unroll\n\t\t\t\t\tfor (int j = 0; j < 4; ++j)\n"
			  "\t\t\t\t\t\t%s[4 * q + j] = __builtin_amdgcn_raw_buffer_load_b32(rs, ((g >> j) & 1u) ? (rr + j) * 4u : OOB, 0, 0);\n\t\t\t\t\t}\n",
			  lo, lo, lo, lo, lo);
		} else if (cl.width == 8) {
			o("\t\t\t\t\tif constexpr (AL) {\n\t\t\t\t\tconst u32x4 v0 = __builtin_amdgcn_raw_buffer_load_b128(rs, g ? rr * 8u : OOB, 0, 0);\n"
			  "\t\t\t\t\tconst u32x4 v1 = __builtin_amdgcn_raw_buffer_load_b128(rs, g ? rr * 8u + 16u : OOB, 0, 0);\n"
			  "\t\t\t\t\t%s[4 * q] = v0.x; %s[4 * q] = v0.y; %s[4 * q + 1] = v0.z; %s[4 * q + 1] = v0.w;\n"
			  "\t\t\t\t\t%s[4 * q + 2] = v1.x; %s[4 * q + 2] = v1.y; %s[4 * q + 3] = v1.z; %s[4 * q + 3] = v1.w;\n"
			  "\t\t\t\t\t} else {\n#pragma unroll\n\t\t\t\t\tfor (int j = 0; j < 4; ++j) {\n"
			  "\t\t\t\t\t\tconst u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, ((g >> j) & 1u) ? (rr + j) * 8u : OOB, 0, 0);\n"
			  "\t\t\t\t\t\t%s[4 * q + j] = v.x; %s[4 * q + j] = v.y;\n\t\t\t\t\t}\n\t\t\t\t\t}\n",
			  lo, hi, lo, hi, lo, hi, lo, hi, lo, hi);
		} else {
			o("\t\t\t\t\tif constexpr (AL) {\n\t\t\t\t\tconst u32 v = __builtin_amdgcn_raw_buffer_load_b32(rs, g ? rr : OOB, 0, 0);\n"
			  "#pragma unroll\n\t\t\t\t\tfor (int j = 0; j < 4; ++j) %s[4 * q + j] = (u32)(i32)(i8)(v >> (8 * j));\n"
			  "\t\t\t\t\t} else {\n#pragma unroll\n\t\t\t\t\tfor (int j = 0; j < 4; ++j)\n"
			  "\t\t\t\t\t\t%s[4 * q + j] = (u32)(i32)(i8)__builtin_amdgcn_raw_buffer_load_b8(rs, ((g >> j) & 1u) ? rr + j : OOB, 0, 0);\n\t\t\t\t\t}\n",
			  lo, lo);
		}
		o("\t\t\t\t}\n\t\t\t}\n");
	};

	/* a held shape's stage(): every tile's filter column(s) into the
	 * chunk's registers, all loads in flight together (rows past a
	 * tablet's end read 0 through the resource's size) */
	if (sh->held) {
		o("template <typename AT>\n__device__ __attribute__((always_inline)) inline void stage(AT &A, const u32 c, u8 *lds, held &H)\n{\n");
		o("\t(void)lds;\n");
		o("%s", pre_fn);
		o("#pragma unroll\n\t\tfor (u32 s = 0; s < S_; ++s) {\n\t\t\tif (s >= nt)\n\t\t\t\tbreak;\n%s", tile_info);
		o("\t\t\t(void)fast;\n\t\t\tconst u32 sel = RMASK;\n");
		for (uint32_t c = 0; c < P->ncol; c++)
			if (is_filter[c]) {
				char tl[48], th[48];
				snprintf(tl, sizeof tl, "H.f%u[s]", c);
				snprintf(th, sizeof th, "H.f%uh[s]", c);
				col_load(c, tl, th);
			}
		o("\t\t}\n\t}\n}\n\n");
	}
	/* a held shape: column c's values of tile s from the chunk's registers */
	auto held_read = [&](uint32_t c, const char *ind) {
		const vg_col &cl = P->col[c];
		o("#pragma unroll\n%sfor (int r = 0; r < R; ++r) s%u[r] = H.f%u[s][r];\n", ind, cl.slot, c);
		if (cl.width == 8)
			o("#pragma unroll\n%sfor (int r = 0; r < R; ++r) s%u[r] = H.f%uh[s][r];\n", ind, cl.slot_hi, c);
	};

	/* phase 1 evaluation */
	o("template <typename AT>\n__device__ __attribute__((always_inline)) inline u32 phase1(AT &A, const u32 c, "
	  "const u64 epoch, u32 (*s_cnt)[4], u8 *lds, u64 &vb_out, held &H)\n{\n");
	o("\t(void)H;\n");
	o("%s", pre_fn);
	if (sh->staged)
		o("\t\twait_vm();          /* the staged filter columns */\n");
	o("\t\tu64 vb = 0;\n");
	/* a held shape indexes its registers by tile: the loop is unrolled */
	if (sh->held)
		o("#pragma unroll\n\t\tfor (u32 s = 0; s < S_; ++s) {\n\t\t\tif (s >= nt)\n\t\t\t\tbreak;\n%s", tile_info);
	else
		o("\t\tfor (u32 s = 0; s < nt; ++s) {\n%s", tile_info);
	o("\t\t\t(void)fast;\n");
	o("\t\t\tu32 live = RMASK;\n\t\t\tif (nrows < TR) {\n\t\t\t\tlive = 0;\n"
	  "#pragma unroll\n\t\t\t\tfor (int q = 0; q < NQ; ++q)\n#pragma unroll\n\t\t\t\t\tfor (int j = 0; j < 4; ++j)\n"
	  "\t\t\t\t\t\tlive |= (wave * (NQ * 256u) + q * 256u + lane * 4u + j < nrows) ? 1u << (4 * q + j) : 0u;\n\t\t\t}\n");
	o("%s", decl.s.c_str());
	if (sh->held) {
		for (uint32_t c = 0; c < P->ncol; c++)
			if (is_filter[c])
				held_read(c, "\t\t\t");
	} else if (sh->staged) {
		for (uint32_t c = 0; c < P->ncol; c++)
			if (is_filter[c])
				lds_read(o, c, "\t\t\t");
	} else {
		/* filter columns straight into their slots, every live row */
		o("\t\t\t{\n\t\t\tconst u32 sel = live;\n");
		for (uint32_t c = 0; c < P->ncol; c++)
			if (is_filter[c])
				col_load(c);
		o("\t\t\t}\n");
	}
	o("\t\t\tu32 valid = live, active = live, w0 = 0, w1 = 0, w2 = 0, w3 = 0;\n");
	gen_uops(o, P, P->nfilter, ~0ull);
	o("\t\t\t(void)active; (void)w0; (void)w1; (void)w2; (void)w3;\n");
	o("\t\t\tvb |= (u64)valid << (s * R);\n");
	o("\t\t\tconst u32 cnt = wsum((u32)__builtin_popcount(valid));\n");
	o("\t\t\tif (lane == 0)\n\t\t\t\ts_cnt[s][wave] = cnt;\n\t\t}\n");
	o("\t\t__syncthreads();\n");
	o("\t\tu32 n = 0;\n\t\tfor (u32 s = 0; s < nt; ++s)\n\t\t\tn += s_cnt[s][0] + s_cnt[s][1] + s_cnt[s][2] + s_cnt[s][3];\n");
	o("\t\tn = uni(n);\n");
	o("\t\tif (tid == 0)\n\t\t\t__hip_atomic_store(&A.status[c], (c == 0 ? ST_PRE : ST_AGG) | (epoch << ST_VBITS) | n,\n"
	  "\t\t\t\t__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n");
	o("\t\tvb_out = vb;\n\t\treturn n;\n\t}\n}\n\n");

	/* phase 2 */
	o("template <typename AT, typename G>\n__device__ __attribute__((always_inline)) inline u32 phase2(AT &A, const u32 c, "
	  "const u64 epoch, u32 (*s_cnt)[4], u64 *s_pre, u8 *lds, u8 *ost, const u64 vb, const u32 n, held &H, "
	  "G grab_next)\n{\n");
	o("\tu32 ovf = 0;\n\t(void)H;\n");
	o("%s", pre_fn);
	o("\t\tu8 *const ostage = ost + wave * %uu;\n", sh->stage_wave_bytes);
	o("\t\t/* phase 2: result columns, tile by tile */\n");
	o("\t\tu64 pre = 0;\n\t\tu32 tdone = 0;\n");
	/* a held shape indexes its registers by tile: the loop is unrolled */
	if (sh->held)
		o("#pragma unroll\n\t\tfor (u32 s = 0; s < S_; ++s) {\n\t\t\tif (s >= nt)\n\t\t\t\tbreak;\n%s", tile_info);
	else
		o("\t\tfor (u32 s = 0; s < nt; ++s) {\n%s", tile_info);
	/* the next chunk is taken only as this one's last tile starts: a chunk
	 * taken is a chunk worked on, so later chunks never wait for one that
	 * sits behind a whole chunk of its workgroup's work */
	o("\t\t\tif (s + 1 == nt && tid == 0)\n\t\t\t\tgrab_next();\n");
	o("\t\t\tconst u32 sel = (u32)(vb >> (s * R)) & RMASK;\n");
	o("%s", decl.s.c_str());
	/* result-phase columns: project / emit columns, and filter columns
	 * that were not staged */
	for (uint32_t c = 0; c < P->ncol; c++)
		if (need2[c] && !(is_filter[c] && (sh->staged || sh->held)))
			col_load(c);
	/* filter columns again, from LDS or the registers that hold them */
	for (uint32_t c = 0; c < P->ncol; c++)
		if (is_filter[c] && need2[c] && sh->staged)
			lds_read(o, c, "\t\t\t");
	for (uint32_t c = 0; c < P->ncol; c++)
		if (is_filter[c] && need2[c] && sh->held)
			held_read(c, "\t\t\t");
	/* the chunk's output offset: wave 0 looks back while the loads fly */
	o("\t\t\tif (s == 0) {\n\t\t\t\tif (wave == 0) {\n\t\t\t\t\tbool gave = false;\n"
	  "\t\t\t\t\tconst u64 p = lookback(A.status, c, n, lane, epoch, A.spin_limit, &gave);\n"
	  "\t\t\t\t\tif (gave) ovf |= 2u;\n\t\t\t\t\tif (lane == 0) *s_pre = p;\n\t\t\t\t}\n"
	  "\t\t\t\t__syncthreads();\n\t\t\t\tpre = *s_pre;\n"
	  "\t\t\t\tif (c == A.nchunks - 1 && tid == 0) A.ctrl->total = pre + n;\n\t\t\t}\n");
	o("\t\t\tu32 valid = sel, active = sel, w0 = 0, w1 = 0, w2 = 0, w3 = 0;\n");
	if (P->wfull)
		gen_uops(o, P, P->nuop, ~0ull);
	else
		gen_uops(o, P, P->nuop, P->wneed);
	o("\t\t\t(void)valid; (void)active; (void)w0; (void)w1; (void)w2; (void)w3;\n");

	/* outputs: per strip, the wave's selected rows go to consecutive output
	 * rows; staged in LDS so that lanes store consecutive addresses */
	o("\t\t\tu32 woff = 0;\n\t\t\tfor (u32 w = 0; w < wave; ++w) woff += s_cnt[s][w];\n");
	o("\t\t\tu64 at = pre + tdone + uni(woff);\n");
	o("#pragma unroll\n\t\t\tfor (int q = 0; q < NQ; ++q) {\n"
	  "\t\t\t\tconst u32 g = (sel >> (4 * q)) & 0xFu;\n\t\t\t\tu32 pos = 0, qn = 0;\n"
	  "#pragma unroll\n\t\t\t\tfor (int j = 0; j < 4; ++j) {\n\t\t\t\t\tconst u64 m = __ballot((g >> j) & 1u);\n"
	  "\t\t\t\t\tpos += rank_below(m);\n\t\t\t\t\tqn += (u32)__builtin_popcountll(m);\n\t\t\t\t}\n"
	  "\t\t\t\tqn = uni(qn);\n");
	o("\t\t\t\tif (qn) {\n");
	/* groups of outputs staged together */
	{
		uint32_t j = 0;
		while (j < P->nout) {
			uint32_t gw = 0, j1 = j;
			std::vector<uint32_t> soff;
			while (j1 < P->nout && gw + P->out[j1].width <= 16u) {
				soff.push_back(gw * 256u);
				gw += P->out[j1].width;
				j1++;
			}
			if (j1 == j) {          /* cannot happen: widths are at most 8 */
				soff.push_back(0);
				j1 = j + 1;
			}
			o("\t\t\t\t\t{\n\t\t\t\t\t\tu32 k = pos;\n#pragma unroll\n\t\t\t\t\t\tfor (int j = 0; j < 4; ++j) {\n"
			  "\t\t\t\t\t\t\tif ((g >> j) & 1u) {\n\t\t\t\t\t\t\t\tconst int r = 4 * q + j;\n\t\t\t\t\t\t\t\t(void)r;\n");
			for (uint32_t jj = j; jj < j1; jj++) {
				const outsrc &sr = os[jj];
				const uint32_t so = soff[jj - j];
				char v[128];
				if (sr.kind == VG_OUT_CONST) {
					snprintf(v, sizeof v, "A.oimm[%u]", jj);
				} else if (sr.kind == VG_OUT_COLUMN) {
					if (sr.width == 8)
						snprintf(v, sizeof v, "(((u64)e%dh[r] << 32) | e%d[r])", sr.lo, sr.lo);
					else
						snprintf(v, sizeof v, "e%d[r]", sr.lo);
				} else {
					if (sr.width == 8)
						snprintf(v, sizeof v, "(((u64)s%d[r] << 32) | s%d[r])", sr.hi, sr.lo);
					else
						snprintf(v, sizeof v, "s%d[r]", sr.lo);
				}
				const char *ty = sr.width == 8 ? "u64" : sr.width == 4 ? "u32" : "u8";
				o("\t\t\t\t\t\t\t\t((%s *)(ostage + %u))[k] = (%s)(%s);\n", ty, so, ty, v);
			}
			o("\t\t\t\t\t\t\t\t++k;\n\t\t\t\t\t\t\t}\n\t\t\t\t\t\t}\n\t\t\t\t\t\twave_sync();\n");
			o("\t\t\t\t\t\tfor (u32 e = lane; e < qn; e += 64u) {\n\t\t\t\t\t\t\tconst u64 gi = at + e;\n"
			  "\t\t\t\t\t\t\tif (gi < A.capacity) {\n");
			for (uint32_t jj = j; jj < j1; jj++) {
				const uint32_t so = soff[jj - j];
				const char *ty = os[jj].width == 8 ? "u64" : os[jj].width == 4 ? "u32" : "u8";
				o("\t\t\t\t\t\t\t\t((%s *)A.out[%u])[gi] = ((const %s *)(ostage + %u))[e];\n", ty, jj, ty, so);
			}
			o("\t\t\t\t\t\t\t} else {\n\t\t\t\t\t\t\t\tovf |= 1u;\n\t\t\t\t\t\t\t}\n\t\t\t\t\t\t}\n\t\t\t\t\t\twave_sync();\n\t\t\t\t\t}\n");
			j = j1;
		}
	}
	o("\t\t\t\t}\n\t\t\t\tat += qn;\n\t\t\t}\n");
	o("\t\t\ttdone += s_cnt[s][0] + s_cnt[s][1] + s_cnt[s][2] + s_cnt[s][3];\n\t\t}\n");
	o("\t}\n\treturn ovf;\n}\n\n");

	/* the three steps in a row, with the output staging after one buffer */
	if (outlined)
		o("template <typename AT, typename G>\n__device__ __attribute__((noinline)) u32 chunk(AT &A, const u32 c, "
		  "const u64 epoch, las u32 (*s_cnt3)[4], las u64 *s_pre3, las u8 *lds3, G grab_next)\n{\n"
		  "\tu32 (*const s_cnt)[4] = (u32 (*)[4])s_cnt3;\n\tu64 *const s_pre = (u64 *)s_pre3;\n"
		  "\tu8 *const lds = (u8 *)lds3;\n");
	else
		o("template <typename AT, typename G>\n__device__ __attribute__((always_inline)) inline u32 chunk(AT &A, "
		  "const u32 c, const u64 epoch, u32 (*s_cnt)[4], u64 *s_pre, u8 *lds, G grab_next)\n{\n");
	o("\theld H;\n\tstage(A, c, lds, H);\n\tu64 vb = 0;\n\tconst u32 n = phase1(A, c, epoch, s_cnt, lds, vb, H);\n"
	  "\treturn phase2(A, c, epoch, s_cnt, s_pre, lds, lds + %uu, vb, n, H, grab_next);\n}\n", sh->stage_off);
	o("} /* namespace %s */\n", ns);
}

static void gen_prelude(out_buf &o)
{
	o.s = "typedef unsigned char uint8_t; typedef unsigned short uint16_t; typedef unsigned int uint32_t;\n"
	      "typedef unsigned long long uint64_t; typedef int int32_t; typedef long long int64_t;\n";
	o.s += k_headers;
	o.s += k_prelude;
}

/* Generate the kernel source of program P at shape sh. */
extern "C" int vgx_source(const vg_prog *P, const vgx_shape *sh, char **out_src)
{
	out_buf o;
	gen_prelude(o);
	gen_program(o, P, sh, "p0");
	if (sh->waves > 0)
		o("\nextern \"C\" __global__ __launch_bounds__(256, %d) void vgx_main(const vgx_args A)\n{\n", sh->waves);
	else
		o("\nextern \"C\" __global__ __launch_bounds__(256) void vgx_main(const vgx_args A)\n{\n");
	o("\textern __shared__ __attribute__((aligned(16))) u8 lds[];\n");
	o("\t__shared__ u32 s_cnt[%u][%u][4];\n\t__shared__ u64 s_pre;\n\t__shared__ u32 s_chunk;\n\t__shared__ u32 s_epoch;\n",
	  sh->pipe ? 2u : 1u, sh->S);
	o("\tconst u32 tid = threadIdx.x;\n");
	o("\tif (tid == 0) {\n"
	  "\t\tconst u64 e = __hip_atomic_load(&A.state[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
	  "\t\twait_vm();   /* the epoch is in hand before the first grab */\n"
	  "\t\ts_epoch = (u32)e;\n\t\ts_chunk = grab(A, e);\n\t}\n\t__syncthreads();\n");
	o("\tconst u64 epoch = s_epoch;\n\tu32 ovf = 0;\n");
	if (sh->pipe) {
		/* Pipelined: a chunk's filter phase (which waits for nothing but
		 * its own loads) runs right after the chunk is taken, and the
		 * previous chunk's look-back and writes come after it, so the
		 * chunks before that one have had a whole filter phase to publish
		 * their counts.  Every chunk taken has its count published
		 * without waiting on another, so no look-back can wait for ever. */
		/* The chunk after next is taken as this chunk's last tile starts
		 * (thread 0's atomic flies under that tile's work): a workgroup
		 * makes exactly one grab past the last chunk. */
		const uint32_t buf = sh->stage_off / 2u;
		if (!sh->held) {
		o("\tp0::held Hd;\n");
		o("\tu32 c0 = uni(s_chunk), c1 = 0xFFFFFFFFu, b = 0, n0 = 0;\n\tu64 vb0 = 0;\n");
		o("\tif (c0 < A.nchunks) {\n\t\tp0::stage(A, c0, lds, Hd);\n"
		  "\t\tn0 = p0::phase1(A, c0, epoch, s_cnt[0], lds, vb0, Hd);\n"
		  "\t\t__syncthreads();\n\t\tif (tid == 0)\n\t\t\ts_chunk = grab(A, epoch);\n\t\t__syncthreads();\n"
		  "\t\tc1 = uni(s_chunk);\n\t}\n");
		o("\twhile (c0 < A.nchunks) {\n"
		  "\t\tu32 n1 = 0;\n\t\tu64 vb1 = 0;\n"
		  "\t\tu8 *const buf1 = lds + (b ^ 1u) * %uu;\n"
		  "\t\tif (c1 < A.nchunks) {\n\t\t\tp0::stage(A, c1, buf1, Hd);\n"
		  "\t\t\tn1 = p0::phase1(A, c1, epoch, s_cnt[b ^ 1u], buf1, vb1, Hd);\n\t\t}\n"
		  "\t\tu32 nxt = 0xFFFFFFFFu;\n"
		  "\t\tconst bool more = c1 < A.nchunks;\n"
		  "\t\tovf |= p0::phase2(A, c0, epoch, s_cnt[b], &s_pre, lds + b * %uu, lds + %uu, vb0, n0, Hd,\n"
		  "\t\t\t[&]() { if (more) nxt = grab(A, epoch); });\n", buf, buf, sh->stage_off);
		o("\t\t__syncthreads();   /* s_chunk, and the buffer the next chunk stages into */\n"
		  "\t\tif (tid == 0)\n\t\t\ts_chunk = nxt;\n\t\t__syncthreads();\n"
		  "\t\tc0 = c1;\n\t\tc1 = uni(s_chunk);\n\t\tn0 = n1;\n\t\tvb0 = vb1;\n\t\tb ^= 1u;\n\t}\n");
		} else {
		/* held: the two chunks' registers take turns, the loop unrolled by
		 * two so that each half names its own */
		o("\tp0::held H0, H1;\n");
		o("\tu32 c0 = uni(s_chunk), c1 = 0xFFFFFFFFu, n0 = 0;\n\tu64 vb0 = 0;\n");
		o("\tif (c0 < A.nchunks) {\n\t\tp0::stage(A, c0, lds, H0);\n"
		  "\t\tn0 = p0::phase1(A, c0, epoch, s_cnt[0], lds, vb0, H0);\n"
		  "\t\t__syncthreads();\n\t\tif (tid == 0)\n\t\t\ts_chunk = grab(A, epoch);\n\t\t__syncthreads();\n"
		  "\t\tc1 = uni(s_chunk);\n\t}\n");
		o("\tfor (;;) {\n");
		for (int half = 0; half < 2; half++) {
			const char *cur = half ? "H1" : "H0", *nx = half ? "H0" : "H1";
			const unsigned bc = half ? 1u : 0u, bn = half ? 0u : 1u;
			o("\t\tif (c0 >= A.nchunks)\n\t\t\tbreak;\n\t\t{\n"
			  "\t\t\tu32 n1 = 0;\n\t\t\tu64 vb1 = 0;\n"
			  "\t\t\tif (c1 < A.nchunks) {\n\t\t\t\tp0::stage(A, c1, lds, %s);\n"
			  "\t\t\t\tn1 = p0::phase1(A, c1, epoch, s_cnt[%u], lds, vb1, %s);\n\t\t\t}\n"
			  "\t\t\tu32 nxt = 0xFFFFFFFFu;\n"
			  "\t\t\tconst bool more = c1 < A.nchunks;\n"
			  "\t\t\tovf |= p0::phase2(A, c0, epoch, s_cnt[%u], &s_pre, lds, lds + %uu, vb0, n0, %s,\n"
			  "\t\t\t\t[&]() { if (more) nxt = grab(A, epoch); });\n"
			  "\t\t\t__syncthreads();\n"
			  "\t\t\tif (tid == 0)\n\t\t\t\ts_chunk = nxt;\n\t\t\t__syncthreads();\n"
			  "\t\t\tc0 = c1;\n\t\t\tc1 = uni(s_chunk);\n\t\t\tn0 = n1;\n\t\t\tvb0 = vb1;\n\t\t}\n",
			  nx, bn, nx, bc, sh->stage_off, cur);
		}
		o("\t}\n");
		}
	} else {
		o("\tfor (;;) {\n\t\tconst u32 c = uni(s_chunk);\n\t\tif (c >= A.nchunks)\n\t\t\tbreak;\n");
		o("\t\tu32 nxt = 0;\n");
		o("\t\tovf |= p0::chunk(A, c, epoch, s_cnt[0], &s_pre, lds, [&]() { nxt = grab(A, epoch); });\n");
		o("\t\t__syncthreads();\n\t\tif (tid == 0)\n\t\t\ts_chunk = nxt;\n\t\t__syncthreads();\n\t}\n");
	}
	o("\tif (ovf)\n\t\tatomicOr(&A.ctrl->overflow, ovf);\n}\n");

	char *m = (char *)malloc(o.s.size() + 1);
	if (!m)
		return -1;
	memcpy(m, o.s.c_str(), o.s.size() + 1);
	*out_src = m;
	return 0;
}

/* ------------------------------------------------------ compile and cache */
struct vgx_kernel {
	hipModule_t mod = nullptr;
	hipFunction_t fn = nullptr;
	hipFunction_t fn2 = nullptr;   /* a batch module's reset kernel */
	vgx_shape sh;
	int per_cu = 0;            /* workgroups per CU at this kernel's resources */
	double compile_ms = 0;
};

/* scratch a lane may use for spills under a shape's launch bound (the
 * benchmark's held filter spills 24 bytes at three waves per SIMD) */
#define VGX_MAX_SPILL_BYTES 64

static std::mutex g_mu;
/* compiled kernels by source, kept for the life of the process (a kernel
 * rebuilt without its launch bound is found under both sources) */
static std::unordered_map<std::string, vgx_kernel *> g_cache;
static std::string g_arch;

static const char *arch_name()
{
	if (g_arch.empty()) {
		int dev = 0;
		hipDeviceProp_t prop;
		if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
			g_arch = prop.gcnArchName;
		else
			g_arch = "gfx950";
	}
	return g_arch.c_str();
}

/* With VG_JIT_DUMP=<dir>: the source and code object of every module
 * compiled, for inspection. */
static void dump(const std::string &src, const std::vector<char> &code)
{
	const char *dir = getenv("VG_JIT_DUMP");
	if (!dir)
		return;
	char path[512];
	const size_t h = std::hash<std::string>()(src);
	snprintf(path, sizeof path, "%s/vgx_%016zx.hip", dir, h);
	if (FILE *f = fopen(path, "w")) {
		fwrite(src.data(), 1, src.size(), f);
		fclose(f);
	}
	snprintf(path, sizeof path, "%s/vgx_%016zx.co", dir, h);
	if (FILE *f = fopen(path, "wb")) {
		fwrite(code.data(), 1, code.size(), f);
		fclose(f);
	}
}

/* Compile `src` for the current device; NULL with a message on failure. */
static vgx_kernel *compile(const std::string &src, const vgx_shape &sh, char *err, size_t errlen,
	const char *fname = "vgx_main", const char *fname2 = nullptr)
{
	double t0 = 0;
	{
		struct timespec ts;
		clock_gettime(CLOCK_MONOTONIC, &ts);
		t0 = ts.tv_sec * 1e3 + ts.tv_nsec / 1e6;
	}
	hiprtcProgram prog;
	if (hiprtcCreateProgram(&prog, src.c_str(), "vgx.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
		snprintf(err, errlen, "hiprtc: cannot create program");
		return nullptr;
	}
	std::string archopt = std::string("--offload-arch=") + arch_name();
	/* the options the library's own kernels are built with: IEEE f32
	 * arithmetic operation by operation (no contraction), no packed fp32
	 * and no v_bitop3 (this compiler has folded three-input mask logic
	 * into v_bitop3 with a wrong truth table: a filter that drops the
	 * rows failing two OR-ed compares dropped those passing the first) */
	const char *opts[] = {archopt.c_str(), "-O3", "-std=c++17", "-ffp-contract=off",
		"-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
		"-Xclang", "-target-feature", "-Xclang", "-bitop3-insts"};
	hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
	if (rc != HIPRTC_SUCCESS) {
		size_t ls = 0;
		hiprtcGetProgramLogSize(prog, &ls);
		std::string log(ls + 1, '\0');
		hiprtcGetProgramLog(prog, &log[0]);
		snprintf(err, errlen, "hiprtc: %s: %.400s", hiprtcGetErrorString(rc), log.c_str());
		hiprtcDestroyProgram(&prog);
		return nullptr;
	}
	size_t cs = 0;
	hiprtcGetCodeSize(prog, &cs);
	std::vector<char> code(cs);
	hiprtcGetCode(prog, code.data());
	hiprtcDestroyProgram(&prog);
	dump(src, code);
	vgx_kernel *k = new vgx_kernel();
	k->sh = sh;
	if (hipModuleLoadData(&k->mod, code.data()) != hipSuccess ||
		hipModuleGetFunction(&k->fn, k->mod, fname) != hipSuccess ||
		(fname2 && hipModuleGetFunction(&k->fn2, k->mod, fname2) != hipSuccess)) {
		(void)hipGetLastError();
		snprintf(err, errlen, "cannot load the compiled kernel");
		if (k->mod)
			(void)hipModuleUnload(k->mod);
		delete k;
		return nullptr;
	}
	int per = 0;
	if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per, k->fn, 256, sh.lds_bytes) != hipSuccess || per < 1) {
		(void)hipGetLastError();
		per = 1;
	}
	k->per_cu = per;
	{
		struct timespec ts;
		clock_gettime(CLOCK_MONOTONIC, &ts);
		k->compile_ms = ts.tv_sec * 1e3 + ts.tv_nsec / 1e6 - t0;
	}
	return k;
}

static int env_int(const char *name, int dflt)
{
	const char *s = getenv(name);
	return s && *s ? atoi(s) : dflt;
}

extern "C" vgx_kernel *vgx_get(const vg_prog *P, int aligned, char *err, size_t errlen)
{
	vgx_shape sh;
	if (vgx_shape_of(P, aligned, env_int("VG_JIT_R", 0), env_int("VG_JIT_S", 0), env_int("VG_JIT_PIPE", 1), &sh,
			err, errlen) != 0)
		return nullptr;
	char *src = nullptr;
	if (vgx_source(P, &sh, &src) != 0) {
		snprintf(err, errlen, "out of memory");
		return nullptr;
	}
	std::string key(src);
	free(src);
	std::lock_guard<std::mutex> g(g_mu);
	auto it = g_cache.find(key);
	if (it != g_cache.end())
		return it->second;
	vgx_kernel *k = compile(key, sh, err, errlen);
	if (k && sh.waves > 0) {
		/* the launch bound may make the compiler spill; a few dwords a
		 * lane cost less than the occupancy they buy, more do not: then
		 * the kernel is built again at the compiler's own occupancy */
		int scratch = 0;
		if (hipFuncGetAttribute(&scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, k->fn) != hipSuccess) {
			(void)hipGetLastError();
			scratch = 0;
		}
		if (scratch > VGX_MAX_SPILL_BYTES) {
			(void)hipModuleUnload(k->mod);
			delete k;
			sh.waves = 0;
			if (vgx_source(P, &sh, &src) != 0) {
				snprintf(err, errlen, "out of memory");
				return nullptr;
			}
			std::string key2(src);
			free(src);
			auto it2 = g_cache.find(key2);
			k = it2 != g_cache.end() ? it2->second : compile(key2, sh, err, errlen);
			if (k)
				g_cache[key2] = k;
		}
	}
	if (k)
		g_cache[key] = k;
	return k;
}

extern "C" const vgx_shape *vgx_kernel_shape(const vgx_kernel *k)
{
	return &k->sh;
}

extern "C" double vgx_kernel_compile_ms(const vgx_kernel *k)
{
	return k->compile_ms;
}

extern "C" unsigned vgx_kernel_grid(const vgx_kernel *k, unsigned nchunks, unsigned cus)
{
	unsigned g = (unsigned)k->per_cu * cus;
	return g < nchunks ? g : nchunks;
}

extern "C" int vgx_launch(const vgx_kernel *k, const vgx_launch_args *a, unsigned grid, hipStream_t stream)
{
	size_t sz = sizeof *a;
	void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, (void *)a, HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz,
		HIP_LAUNCH_PARAM_END};
	if (grid == 0)
		return 0;
	const hipError_t e = hipModuleLaunchKernel(k->fn, grid, 1, 1, 256, 1, 1, k->sh.lds_bytes, stream,
		nullptr, cfg);
	return e == hipSuccess ? 0 : -2;
}

/* ------------------------------------------------------------- batches */
/* A batch's programs are cut into chunks of 8K rows (their filter
 * staging half the one-query kernel's 16K-row chunks): three workgroups
 * per CU fit the LDS instead of two, 1.34 against 1.41 ms per run of the
 * 256-program batch.  VG_JIT_BATCH_SMALL=0 keeps 16K-row chunks. */
extern "C" int vgxb_shape_mode(void)
{
	const char *e = getenv("VG_JIT_BATCH_SMALL");
	return e && e[0] == '0' ? 0 : 2;
}

/* Source of a batch module: every distinct program of the batch once (as
 * namespace p<k>), a reset kernel and the persistent kernel.  fn[i] is
 * query i's program; lds: the dynamic LDS the kernel takes. */
static std::string vgxb_text(const vg_prog *const *P, const int *aligned, unsigned n, uint32_t *fn,
	uint32_t *lds, char *err, size_t errlen)
{
	out_buf o;
	gen_prelude(o);
	std::unordered_map<std::string, uint32_t> seen;
	uint32_t nk = 0, maxlds = 0, maxs = 1;
	for (unsigned i = 0; i < n; i++) {
		vgx_shape sh;
		if (vgx_shape_of(P[i], aligned[i], 0, 0, vgxb_shape_mode(), &sh, err, errlen) != 0)
			return std::string();
		out_buf body;
		gen_program(body, P[i], &sh, "pK", true);
		auto it = seen.find(body.s);
		if (it != seen.end()) {
			fn[i] = it->second;
			continue;
		}
		char ns[32];
		snprintf(ns, sizeof ns, "p%u", nk);
		gen_program(o, P[i], &sh, ns, true);
		seen[body.s] = nk;
		fn[i] = nk++;
		maxlds = sh.lds_bytes > maxlds ? sh.lds_bytes : maxlds;
		maxs = sh.S > maxs ? sh.S : maxs;
	}
	*lds = maxlds;
	o(R"VGXB(
typedef const __attribute__((address_space(4))) vgxb_args cvgxb_args;

/* the next work item; the grab completing the launch's nitems + grid
 * grabs sets the counter back and moves the epoch on */
__device__ __forceinline__ u32 grab_item(u64 *state, u32 nitems, u64 epoch)
{
	const u64 g = atomicAdd(&state[0], 1ull);
	if (g == (u64)nitems + gridDim.x - 1) {
		__hip_atomic_store(&state[0], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		__hip_atomic_store(&state[1], (epoch + 1) & ST_EMASK, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}
	return g < nitems ? (u32)g : 0xFFFFFFFFu;
}

/* thread 0's call for the next item, which it leaves in LDS: plain values
 * (the chunk functions are not inlined, and a closure over the kernel's
 * locals would put them in scratch) */
struct next_item {
	u64 *state;
	u32 nitems;
	u64 epoch;
	u32 *slot;
	__device__ void operator()() const { *slot = grab_item(state, nitems, epoch); }
};

/* each query's control block, chunk count and outcome cleared; the start stamped */
extern "C" __global__ __launch_bounds__(256) void vgxb_reset(const vgxb_args B)
{
	const u32 i = blockIdx.x * 256u + threadIdx.x;
	if (i < B.nquery) {
		vg_ctrl *c = B.qa[i].ctrl;
		c->total = 0;
		c->overflow = 0;
		B.done[i] = 0;
		vg_bresult r;
		r.total = 0; r.t_done = 0; r.t_start = 0; r.flags = 0; r.done = 0;
		B.hres[i] = r;
	}
	if (i == 0)
		B.state[2] = wall_clock64();
}

)VGXB");
	/* Occupancy bound for the chunk functions' registers (they are not
	 * compiled under the kernel's own occupancy target): three workgroups
	 * per CU, or what the LDS allows when that is fewer. */
	const uint32_t by_lds = 163840u / (maxlds + 1024u);
	const uint32_t waves = by_lds < 1 ? 1 : by_lds > 3 ? 3 : by_lds;
	o("extern \"C\" __global__ __launch_bounds__(256, %u) void vgxb_main(const vgxb_args B)\n{\n"
	  "\textern __shared__ __attribute__((aligned(16))) u8 lds[];\n", waves);
	o("\t__shared__ u32 s_cnt[%u][4];\n", maxs);
	o(R"VGXB(	__shared__ u64 s_pre;
	__shared__ u32 s_item, s_next, s_epoch, s_ovf;
	const u32 tid = threadIdx.x;
	if (tid == 0) {
		const u64 e = __hip_atomic_load(&B.state[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		wait_vm();   /* the epoch is in hand before the first grab */
		s_epoch = (u32)e;
		s_ovf = 0;
		s_item = grab_item(B.state, B.nitems, e);
	}
	__syncthreads();
	const u64 epoch = s_epoch;
	for (;;) {
		const u32 i = uni(s_item);
		if (i >= B.nitems)
			break;
		const vgxb_item it = {ldc(&B.items[i].q), ldc(&B.items[i].c)};
		const u32 q = uni(it.q), c = uni(it.c);
		cvgx_args &A = *(cvgx_args *)(B.qa + q);
		const next_item next = {B.state, B.nitems, epoch, &s_next};
		u32 ovf = 0;
		switch (ldc(B.fn + q)) {
)VGXB");
	for (uint32_t k = 0; k < nk; k++)
		o("\t\tcase %u: ovf = p%u::chunk(A, c, epoch, (las u32 (*)[4])s_cnt, (las u64 *)&s_pre, (las u8 *)lds, next); "
		  "break;\n", k, k);
	o(R"VGXB(		}
		if (ovf)
			atomicOr(&s_ovf, ovf);
		__syncthreads();
		if (tid == 0) {
			const u32 f = s_ovf;
			s_ovf = 0;
			if (f)
				atomicOr(&A.ctrl->overflow, f);
			/* the chunk that wrote the query's total, and any that set flags,
			 * publish before counting themselves done; the last one done
			 * reads the query's outcome after them */
			const bool pub = c + 1 == A.nchunks || f != 0;
			const u32 d = pub ? __hip_atomic_fetch_add(&B.done[q], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT)
			                  : __hip_atomic_fetch_add(&B.done[q], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			if (d + 1 == A.nchunks) {
				__atomic_thread_fence(__ATOMIC_ACQUIRE);
				vg_bresult r;
				r.total = __hip_atomic_load(&A.ctrl->total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				r.flags = __hip_atomic_load(&A.ctrl->overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				r.t_done = wall_clock64();
				r.t_start = B.state[2];
				r.done = 0;
				B.hres[q] = r;
				/* the fields, then the flag a polling host waits for */
				__threadfence_system();
				__hip_atomic_store(&B.hres[q].done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
			}
			s_item = s_next;
		}
		__syncthreads();
	}
}
)VGXB");
	return o.s;
}

extern "C" vgx_kernel *vgxb_get(const vg_prog *const *P, const int *aligned, unsigned n, uint32_t *fn,
	char *err, size_t errlen)
{
	uint32_t lds = 0;
	const std::string src = vgxb_text(P, aligned, n, fn, &lds, err, errlen);
	if (src.empty())
		return nullptr;
	std::lock_guard<std::mutex> g(g_mu);
	auto it = g_cache.find(src);
	if (it != g_cache.end())
		return it->second;
	vgx_shape sh;
	memset(&sh, 0, sizeof sh);
	sh.lds_bytes = lds;
	vgx_kernel *k = compile(src, sh, err, errlen, "vgxb_main", "vgxb_reset");
	if (k)
		g_cache[src] = k;
	return k;
}

extern "C" unsigned vgxb_grid(const vgx_kernel *k, unsigned nitems, unsigned cus)
{
	return vgx_kernel_grid(k, nitems, cus);
}

/* the reset, then the persistent launch, on `stream` */
extern "C" int vgxb_launch(const vgx_kernel *k, const vgxb_args *a, unsigned grid, hipStream_t stream)
{
	size_t sz = sizeof *a;
	void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, (void *)a, HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz,
		HIP_LAUNCH_PARAM_END};
	const unsigned rgrid = (a->nquery + 255) / 256;
	if (hipModuleLaunchKernel(k->fn2, rgrid ? rgrid : 1, 1, 1, 256, 1, 1, 0, stream, nullptr, cfg) != hipSuccess)
		return -2;
	if (grid == 0)
		return 0;
	return hipModuleLaunchKernel(k->fn, grid, 1, 1, 256, 1, 1, k->sh.lds_bytes, stream, nullptr, cfg) == hipSuccess
		? 0 : -2;
}

/* ------------------------------------------------- inspection (no device) */
extern "C" int vg_decode(const virg_op *stmt, unsigned num_ops, unsigned pc,
	const virg_tablet_meta *tab, const virg_tablet_meta *res, vg_prog *P, char *err, size_t errlen);

/* Decode a program and generate its specialised kernel's source into buf
 * (truncated to n bytes); returns the source length, or -1 with a message. */
extern "C" long vg_jit_source(const virg_op *ops, unsigned num_ops, unsigned pc,
	const virg_tablet_meta *tab, char *buf, size_t n, char *err, size_t errlen)
{
	vg_prog P;
	if (vg_decode(ops, num_ops, pc, tab, nullptr, &P, err, errlen) != VIRG_SUCCESS)
		return -1;
	vgx_shape sh;
	if (vgx_shape_of(&P, 1, env_int("VG_JIT_R", 0), env_int("VG_JIT_S", 0), env_int("VG_JIT_PIPE", 1), &sh, err,
			errlen) != 0)
		return -1;
	char *src = nullptr;
	if (vgx_source(&P, &sh, &src) != 0)
		return -1;
	const size_t len = strlen(src);
	if (buf && n) {
		const size_t m = len < n - 1 ? len : n - 1;
		memcpy(buf, src, m);
		buf[m] = '\0';
	}
	free(src);
	return (long)len;
}

/* Immediates of a decoded program as the specialised kernel reads them. */
extern "C" void vgx_immediates(const vg_prog *P, vgx_launch_args *a)
{
	for (uint32_t k = 0; k < VG_MAX_UOPS; k++) {
		a->imm[k][0] = k < P->nuop ? P->uop[k].ia : 0;
		a->imm[k][1] = k < P->nuop ? P->uop[k].ib : 0;
	}
	for (uint32_t j = 0; j < VG_MAX_OUT; j++)
		a->oimm[j] = j < P->nout ? P->out[j].imm : 0;
}

/* Everything a launch of a program's specialised kernel takes besides its
 * tablets (no device needed): the kernel's shape for tablets like `tab`
 * (aligned: the 16-byte aligned variant), the immediates (a->imm, a->oimm),
 * the result columns' widths (out_width[VG_MAX_OUT]) and the source (into
 * buf, n bytes).  Returns the source length, or -1 with a message. */
extern "C" long vg_jit_program(const virg_op *ops, unsigned num_ops, unsigned pc, const virg_tablet_meta *tab,
	int aligned, vgx_shape *sh, vgx_launch_args *a, unsigned *nout, unsigned *out_width, char *buf, size_t n,
	char *err, size_t errlen)
{
	vg_prog P;
	if (vg_decode(ops, num_ops, pc, tab, nullptr, &P, err, errlen) != VIRG_SUCCESS)
		return -1;
	if (vgx_shape_of(&P, aligned, env_int("VG_JIT_R", 0), env_int("VG_JIT_S", 0), env_int("VG_JIT_PIPE", 1), sh,
			err, errlen) != 0)
		return -1;
	memset(a, 0, sizeof *a);
	vgx_immediates(&P, a);
	*nout = P.nout;
	for (unsigned j = 0; j < VG_MAX_OUT; j++)
		out_width[j] = j < P.nout ? P.out[j].width : 0;
	char *src = nullptr;
	if (vgx_source(&P, sh, &src) != 0)
		return -1;
	const size_t len = strlen(src);
	if (buf && n) {
		const size_t m = len < n - 1 ? len : n - 1;
		memcpy(buf, src, m);
		buf[m] = '\0';
	}
	free(src);
	return (long)len;
}

/* hiprtc compile of src for gfx950, not loaded: 1, or 0 with the log */
static int rtc_check(const char *src, double *ms, char *err, size_t errlen)
{
	struct timespec t0, t1;
	clock_gettime(CLOCK_MONOTONIC, &t0);
	hiprtcProgram prog;
	int ok = 0;
	if (hiprtcCreateProgram(&prog, src, "vgx.hip", 0, nullptr, nullptr) == HIPRTC_SUCCESS) {
		const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
			"-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
			"-Xclang", "-target-feature", "-Xclang", "-bitop3-insts"};
		const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
		if (rc == HIPRTC_SUCCESS) {
			ok = 1;
			size_t cs = 0;
			if (getenv("VG_JIT_DUMP") && hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS) {
				std::vector<char> code(cs);
				if (hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS)
					dump(src, code);
			}
		} else {
			size_t ls = 0;
			hiprtcGetProgramLogSize(prog, &ls);
			std::string log(ls + 1, '\0');
			hiprtcGetProgramLog(prog, &log[0]);
			snprintf(err, errlen, "%s: %.900s", hiprtcGetErrorString(rc), log.c_str());
		}
		hiprtcDestroyProgram(&prog);
	} else {
		snprintf(err, errlen, "hiprtc: cannot create program");
	}
	clock_gettime(CLOCK_MONOTONIC, &t1);
	if (ms)
		*ms = (t1.tv_sec - t0.tv_sec) * 1e3 + (t1.tv_nsec - t0.tv_nsec) / 1e6;
	return ok;
}

/* Compile a program's specialised kernel for gfx950 with hiprtc without
 * loading it (no device needed): 1 on success, 0 with the compiler's
 * message.  *ms: compile time. */
extern "C" int vg_jit_compile_check(const virg_op *ops, unsigned num_ops, unsigned pc,
	const virg_tablet_meta *tab, double *ms, char *err, size_t errlen)
{
	vg_prog P;
	if (vg_decode(ops, num_ops, pc, tab, nullptr, &P, err, errlen) != VIRG_SUCCESS)
		return 0;
	vgx_shape sh;
	if (vgx_shape_of(&P, 1, env_int("VG_JIT_R", 0), env_int("VG_JIT_S", 0), env_int("VG_JIT_PIPE", 1), &sh, err,
			errlen) != 0)
		return 0;
	char *src = nullptr;
	if (vgx_source(&P, &sh, &src) != 0)
		return 0;
	const int ok = rtc_check(src, ms, err, errlen);
	free(src);
	return ok;
}

/* The batch module of n programs (program i: ops[i], its Parallel section
 * at pc[i], over tablets like tab[i]; aligned[i] selects the aligned
 * variant): fn[i] receives the program's index in the module; the module
 * is compiled with hiprtc, not loaded, when `compile` is set.  Returns the
 * number of distinct programs, or 0 with a message; *ms: compile time. */
extern "C" unsigned vg_jitb_check(const virg_op *const *ops, const unsigned *num_ops, const unsigned *pc,
	const virg_tablet_meta *const *tab, const int *aligned, unsigned n, uint32_t *fn, int compile,
	double *ms, char *err, size_t errlen)
{
	std::vector<vg_prog> P(n);
	std::vector<const vg_prog *> pp(n);
	for (unsigned i = 0; i < n; i++) {
		if (vg_decode(ops[i], num_ops[i], pc[i], tab[i], nullptr, &P[i], err, errlen) != VIRG_SUCCESS)
			return 0;
		pp[i] = &P[i];
	}
	uint32_t lds = 0;
	const std::string src = vgxb_text(pp.data(), aligned, n, fn, &lds, err, errlen);
	if (src.empty())
		return 0;
	if (compile && !rtc_check(src.c_str(), ms, err, errlen))
		return 0;
	uint32_t k = 0;
	for (unsigned i = 0; i < n; i++)
		k = fn[i] + 1 > k ? fn[i] + 1 : k;
	return k;
}
