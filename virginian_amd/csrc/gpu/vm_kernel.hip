/*
 * gfx950 kernels of the opcode VM (the data-parallel section of a query:
 * reference src/vm/vm_gpu.cu:1152-1225 on the GPU, src/vm/virginia.c:228-891
 * on the CPU).
 *
 * One workgroup processes one tile of T = NT*R consecutive rows of one
 * tablet; lane l of wave w owns rows (w*64+l)*R .. +R-1, so every column
 * load is one contiguous R*width-byte vector per lane and a wave reads
 * 64*R*width contiguous bytes of HBM.  Tiles are handed out in order by an
 * atomic counter, which is what makes the ordered-output look-back below
 * safe: a tile only waits for tiles taken before it by running workgroups.
 *
 * The VM registers live in LDS as [slot][row of the tile]; a lane's R rows
 * of a slot are contiguous, so reading an operand for all of them is one
 * ds_read_b128 (R = 4) at a wave-uniform slot base.  Micro-op fields are
 * read with scalar loads from the kernel arguments, so dispatch is a
 * scalar branch.
 *
 * Per tile:
 *   1. filter loads: every column read by a validity-changing micro-op is
 *      loaded for all rows (the reference copies whole blocks the same way);
 *   2. filter micro-ops [0, nfilter) with a per-row resume index and
 *      validity flag;
 *   3. projection: only in waves that still hold a valid row, load the
 *      remaining columns for lanes holding one and run [nfilter, nuop);
 *   4. ordered compaction: ballot + mbcnt rank each valid row inside its
 *      wave, an LDS scan over waves inside the tile, and a decoupled
 *      look-back over per-tile status words gives the tile's global offset;
 *      each valid row writes its row number to srcrow[rank];
 *   5. for every result column, consecutive lanes gather srcrow[i] and
 *      store output element base+i: coalesced, in row order.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "vg_prog.h"

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

/* bounded waits: a tile gives up on its predecessors after this many polls
 * (about a second) and reports an error instead of hanging the device */
#define VG_SPIN_LIMIT (1u << 23)

__device__ __forceinline__ uint32_t uni(uint32_t x)
{
	return __builtin_amdgcn_readfirstlane(x);
}

__device__ __forceinline__ uint32_t rank_below(uint64_t m)
{
	return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
		__builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* ------------------------------------------------------- LDS row vectors */
/* R consecutive 32-bit values of one slot for this lane */
template <int R>
struct rv {
	uint32_t v[R];
};

template <int R>
__device__ __forceinline__ rv<R> lds_get(const uint32_t *p)
{
	rv<R> out;
	if constexpr (R % 4 == 0) {
#pragma unroll
		for (int q = 0; q < R / 4; ++q) {
			const u32x4 x = ((const u32x4 *)p)[q];
			out.v[4 * q] = x.x; out.v[4 * q + 1] = x.y; out.v[4 * q + 2] = x.z; out.v[4 * q + 3] = x.w;
		}
	} else if constexpr (R == 2) {
		const u32x2 x = *(const u32x2 *)p;
		out.v[0] = x.x; out.v[1] = x.y;
	} else {
		out.v[0] = p[0];
	}
	return out;
}

template <int R>
__device__ __forceinline__ void lds_put(uint32_t *p, const rv<R> &in)
{
	if constexpr (R % 4 == 0) {
#pragma unroll
		for (int q = 0; q < R / 4; ++q)
			((u32x4 *)p)[q] = u32x4{in.v[4 * q], in.v[4 * q + 1], in.v[4 * q + 2], in.v[4 * q + 3]};
	} else if constexpr (R == 2) {
		*(u32x2 *)p = u32x2{in.v[0], in.v[1]};
	} else {
		p[0] = in.v[0];
	}
}

/* ---------------------------------------------------------------- values */
/* value-type traits: bit patterns <-> values and the reference's
 * arithmetic (int32/int64 wrap, char stored back to 8 bits, IEEE float) */
template <typename T> struct vt;
template <> struct vt<int32_t> {
	static constexpr bool wide = false;
	static __device__ __forceinline__ int32_t from(uint32_t lo, uint32_t) { return (int32_t)lo; }
	static __device__ __forceinline__ uint32_t lo(int32_t v) { return (uint32_t)v; }
	static __device__ __forceinline__ uint32_t hi(int32_t) { return 0; }
	static __device__ __forceinline__ int32_t add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
	static __device__ __forceinline__ int32_t sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
	static __device__ __forceinline__ int32_t mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
	static __device__ __forceinline__ int32_t div(int32_t a, int32_t b) { return b == 0 ? 0 : (b == -1 ? (int32_t)(0u - (uint32_t)a) : a / b); }
};
template <> struct vt<int64_t> {
	static constexpr bool wide = true;
	static __device__ __forceinline__ int64_t from(uint32_t lo, uint32_t hi) { return (int64_t)(((uint64_t)hi << 32) | lo); }
	static __device__ __forceinline__ uint32_t lo(int64_t v) { return (uint32_t)(uint64_t)v; }
	static __device__ __forceinline__ uint32_t hi(int64_t v) { return (uint32_t)((uint64_t)v >> 32); }
	static __device__ __forceinline__ int64_t add(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
	static __device__ __forceinline__ int64_t sub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
	static __device__ __forceinline__ int64_t mul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
	static __device__ __forceinline__ int64_t div(int64_t a, int64_t b) { return b == 0 ? 0 : (b == -1 ? (int64_t)(0ull - (uint64_t)a) : a / b); }
};
template <> struct vt<float> {
	static constexpr bool wide = false;
	static __device__ __forceinline__ float from(uint32_t lo, uint32_t) { return __uint_as_float(lo); }
	static __device__ __forceinline__ uint32_t lo(float v) { return __float_as_uint(v); }
	static __device__ __forceinline__ uint32_t hi(float) { return 0; }
	static __device__ __forceinline__ float add(float a, float b) { return a + b; }
	static __device__ __forceinline__ float sub(float a, float b) { return a - b; }
	static __device__ __forceinline__ float mul(float a, float b) { return a * b; }
	static __device__ __forceinline__ float div(float a, float b) { return a / b; }
};
template <> struct vt<double> {
	static constexpr bool wide = true;
	static __device__ __forceinline__ double from(uint32_t lo, uint32_t hi) { return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo)); }
	static __device__ __forceinline__ uint32_t lo(double v) { return (uint32_t)(uint64_t)__double_as_longlong(v); }
	static __device__ __forceinline__ uint32_t hi(double v) { return (uint32_t)((uint64_t)__double_as_longlong(v) >> 32); }
	static __device__ __forceinline__ double add(double a, double b) { return a + b; }
	static __device__ __forceinline__ double sub(double a, double b) { return a - b; }
	static __device__ __forceinline__ double mul(double a, double b) { return a * b; }
	static __device__ __forceinline__ double div(double a, double b) { return a / b; }
};
/* CHAR: kept sign-extended in a 32-bit slot; arithmetic in int, results
 * truncated to 8 bits as the reference's char registers are */
struct chr { int32_t v; };
template <> struct vt<chr> {
	static constexpr bool wide = false;
	static __device__ __forceinline__ chr from(uint32_t lo, uint32_t) { return chr{(int32_t)lo}; }
	static __device__ __forceinline__ uint32_t lo(chr v) { return (uint32_t)(int32_t)(int8_t)v.v; }
	static __device__ __forceinline__ uint32_t hi(chr) { return 0; }
	static __device__ __forceinline__ chr add(chr a, chr b) { return chr{a.v + b.v}; }
	static __device__ __forceinline__ chr sub(chr a, chr b) { return chr{a.v - b.v}; }
	static __device__ __forceinline__ chr mul(chr a, chr b) { return chr{a.v * b.v}; }
	static __device__ __forceinline__ chr div(chr a, chr b) { return chr{b.v == 0 ? 0 : a.v / b.v}; }
};
__device__ __forceinline__ bool operator<(chr a, chr b) { return a.v < b.v; }
__device__ __forceinline__ bool operator<=(chr a, chr b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>(chr a, chr b) { return a.v > b.v; }
__device__ __forceinline__ bool operator>=(chr a, chr b) { return a.v >= b.v; }
__device__ __forceinline__ bool operator==(chr a, chr b) { return a.v == b.v; }
__device__ __forceinline__ bool operator!=(chr a, chr b) { return a.v != b.v; }
__device__ __forceinline__ bool nz(chr a) { return a.v != 0; }
template <typename T> __device__ __forceinline__ bool nz(T a) { return a != (T)0; }

/* operand of a micro-op: R typed values for this lane */
template <typename T, int R, uint32_t TR>
__device__ __forceinline__ void operand(const uint32_t *regs, bool imm, uint64_t ib, uint32_t lo,
	uint32_t hi, uint32_t lrow, T (&x)[R])
{
	if (imm) {
		const T c = vt<T>::from((uint32_t)ib, (uint32_t)(ib >> 32));
#pragma unroll
		for (int r = 0; r < R; ++r)
			x[r] = c;
		return;
	}
	const rv<R> l = lds_get<R>(regs + lo * TR + lrow);
	if constexpr (vt<T>::wide) {
		const rv<R> h = lds_get<R>(regs + hi * TR + lrow);
#pragma unroll
		for (int r = 0; r < R; ++r)
			x[r] = vt<T>::from(l.v[r], h.v[r]);
	} else {
#pragma unroll
		for (int r = 0; r < R; ++r)
			x[r] = vt<T>::from(l.v[r], 0);
	}
}

/* write R typed values to a slot; masked writes keep inactive rows */
template <typename T, int R, uint32_t TR>
__device__ __forceinline__ void result_put(uint32_t *regs, uint32_t lo, uint32_t hi, uint32_t lrow,
	const T (&x)[R], bool masked, uint32_t k, const uint32_t (&resume)[R])
{
	rv<R> l, h;
	if (masked) {
		l = lds_get<R>(regs + lo * TR + lrow);
		if constexpr (vt<T>::wide)
			h = lds_get<R>(regs + hi * TR + lrow);
	}
#pragma unroll
	for (int r = 0; r < R; ++r) {
		if (!masked || resume[r] <= k) {
			l.v[r] = vt<T>::lo(x[r]);
			if constexpr (vt<T>::wide)
				h.v[r] = vt<T>::hi(x[r]);
		}
	}
	lds_put<R>(regs + lo * TR + lrow, l);
	if constexpr (vt<T>::wide)
		lds_put<R>(regs + hi * TR + lrow, h);
}

/* ------------------------------------------------------------- micro-ops */
template <typename T, int R, uint32_t TR>
__device__ __forceinline__ void exec_typed(const vg_uop &u, uint32_t k, uint32_t *regs,
	uint32_t lrow, uint32_t (&resume)[R], bool (&valid)[R])
{
	T a[R], b[R], x[R];
	operand<T, R, TR>(regs, u.flags & VG_F_AIMM, u.ia, u.a, u.a_hi, lrow, a);
	const uint32_t kind = u.kind;
	if (kind <= VG_K_NOT) {
		if (kind != VG_K_NOT)
			operand<T, R, TR>(regs, u.flags & VG_F_BIMM, u.ib, u.b, u.b_hi, lrow, b);
		const bool keep = u.flags & VG_F_KEEP;
#pragma unroll
		for (int r = 0; r < R; ++r) {
			bool c;
			switch (kind) {
			case VG_K_LT: c = a[r] < b[r]; break;
			case VG_K_LE: c = a[r] <= b[r]; break;
			case VG_K_GT: c = a[r] > b[r]; break;
			case VG_K_GE: c = a[r] >= b[r]; break;
			case VG_K_EQ: c = a[r] == b[r]; break;
			case VG_K_NE: c = a[r] != b[r]; break;
			case VG_K_AND: c = nz(a[r]) && nz(b[r]); break;
			case VG_K_OR: c = nz(a[r]) || nz(b[r]); break;
			default: c = !nz(a[r]); break;
			}
			if (resume[r] <= k && c) {
				valid[r] = valid[r] && keep;
				resume[r] = u.target;
			}
		}
		return;
	}
	if (kind != VG_K_MOV)
		operand<T, R, TR>(regs, u.flags & VG_F_BIMM, u.ib, u.b, u.b_hi, lrow, b);
	switch (kind) {
	case VG_K_ADD:
#pragma unroll
		for (int r = 0; r < R; ++r) x[r] = vt<T>::add(a[r], b[r]);
		break;
	case VG_K_SUB:
#pragma unroll
		for (int r = 0; r < R; ++r) x[r] = vt<T>::sub(a[r], b[r]);
		break;
	case VG_K_MUL:
#pragma unroll
		for (int r = 0; r < R; ++r) x[r] = vt<T>::mul(a[r], b[r]);
		break;
	case VG_K_DIV:
#pragma unroll
		for (int r = 0; r < R; ++r) x[r] = vt<T>::div(a[r], b[r]);
		break;
	default:
#pragma unroll
		for (int r = 0; r < R; ++r) x[r] = a[r];
		break;
	}
	result_put<T, R, TR>(regs, u.dst, u.dst_hi, lrow, x, u.flags & VG_F_MASKED, k, resume);
}

template <typename S, typename D>
__device__ __forceinline__ D convert(S x) { return (D)x; }
template <> __device__ __forceinline__ chr convert<chr, chr>(chr x) { return x; }
template <> __device__ __forceinline__ int32_t convert<chr, int32_t>(chr x) { return x.v; }
template <> __device__ __forceinline__ int64_t convert<chr, int64_t>(chr x) { return x.v; }
template <> __device__ __forceinline__ float convert<chr, float>(chr x) { return (float)x.v; }
template <> __device__ __forceinline__ double convert<chr, double>(chr x) { return (double)x.v; }
template <> __device__ __forceinline__ chr convert<int32_t, chr>(int32_t x) { return chr{(int32_t)(int8_t)x}; }
template <> __device__ __forceinline__ chr convert<int64_t, chr>(int64_t x) { return chr{(int32_t)(int8_t)x}; }
template <> __device__ __forceinline__ chr convert<float, chr>(float x) { return chr{(int32_t)(int8_t)(int32_t)x}; }
template <> __device__ __forceinline__ chr convert<double, chr>(double x) { return chr{(int32_t)(int8_t)(int32_t)x}; }

template <typename S, typename D, int R, uint32_t TR>
__device__ __forceinline__ void cast_rows(const vg_uop &u, uint32_t k, uint32_t *regs, uint32_t lrow,
	const uint32_t (&resume)[R])
{
	S a[R];
	D x[R];
	operand<S, R, TR>(regs, false, 0, u.a, u.a_hi, lrow, a);
#pragma unroll
	for (int r = 0; r < R; ++r)
		x[r] = convert<S, D>(a[r]);
	result_put<D, R, TR>(regs, u.dst, u.dst_hi, lrow, x, true, k, resume);
}

template <typename S, int R, uint32_t TR>
__device__ __forceinline__ void cast_from(const vg_uop &u, uint32_t k, uint32_t *regs, uint32_t lrow,
	const uint32_t (&resume)[R])
{
	switch (u.b) {
	case VG_T_INT: cast_rows<S, int32_t, R, TR>(u, k, regs, lrow, resume); break;
	case VG_T_INT64: cast_rows<S, int64_t, R, TR>(u, k, regs, lrow, resume); break;
	case VG_T_FLOAT: cast_rows<S, float, R, TR>(u, k, regs, lrow, resume); break;
	case VG_T_DOUBLE: cast_rows<S, double, R, TR>(u, k, regs, lrow, resume); break;
	case VG_T_CHAR: cast_rows<S, chr, R, TR>(u, k, regs, lrow, resume); break;
	}
}

template <int R, uint32_t TR>
__device__ __forceinline__ void run_uops(const vg_prog &P, uint32_t kbeg, uint32_t kend,
	uint32_t *regs, uint32_t lrow, uint32_t (&resume)[R], bool (&valid)[R])
{
	for (uint32_t k = kbeg; k < kend; ++k) {
		const vg_uop &u = P.uop[k];
		if (u.kind == VG_K_INVALID) {
#pragma unroll
			for (int r = 0; r < R; ++r)
				if (resume[r] <= k)
					valid[r] = false;
			continue;
		}
		if (u.kind == VG_K_CAST) {
			switch (u.type) {
			case VG_T_INT: cast_from<int32_t, R, TR>(u, k, regs, lrow, resume); break;
			case VG_T_INT64: cast_from<int64_t, R, TR>(u, k, regs, lrow, resume); break;
			case VG_T_FLOAT: cast_from<float, R, TR>(u, k, regs, lrow, resume); break;
			case VG_T_DOUBLE: cast_from<double, R, TR>(u, k, regs, lrow, resume); break;
			case VG_T_CHAR: cast_from<chr, R, TR>(u, k, regs, lrow, resume); break;
			}
			continue;
		}
		switch (u.type) {
		case VG_T_INT: exec_typed<int32_t, R, TR>(u, k, regs, lrow, resume, valid); break;
		case VG_T_INT64: exec_typed<int64_t, R, TR>(u, k, regs, lrow, resume, valid); break;
		case VG_T_FLOAT: exec_typed<float, R, TR>(u, k, regs, lrow, resume, valid); break;
		case VG_T_DOUBLE: exec_typed<double, R, TR>(u, k, regs, lrow, resume, valid); break;
		case VG_T_CHAR: exec_typed<chr, R, TR>(u, k, regs, lrow, resume, valid); break;
		}
	}
}

/* ---------------------------------------------------------- column loads */
/* Rows [lrow, lrow+R) of column c into its LDS slot(s).  The vector path
 * needs the lane's whole group in range and a 16-byte aligned column;
 * otherwise rows are loaded one by one (tails, unaligned tablets). */
template <int R, uint32_t TR>
__device__ __forceinline__ void load_col(const vg_col &c, const char *base, uint32_t lrow,
	uint32_t nrows, bool aligned, uint32_t *regs)
{
	const bool full = aligned && lrow + R <= nrows;
	rv<R> l, h;
#pragma unroll
	for (int r = 0; r < R; ++r)
		l.v[r] = h.v[r] = 0;
	if (c.width == 4) {
		const uint32_t *p = (const uint32_t *)base + lrow;
		if (full) {
			if constexpr (R % 4 == 0) {
#pragma unroll
				for (int q = 0; q < R / 4; ++q) {
					const u32x4 v = __builtin_nontemporal_load((const u32x4 *)p + q);
					l.v[4 * q] = v.x; l.v[4 * q + 1] = v.y; l.v[4 * q + 2] = v.z; l.v[4 * q + 3] = v.w;
				}
			} else if constexpr (R == 2) {
				const u32x2 v = __builtin_nontemporal_load((const u32x2 *)p);
				l.v[0] = v.x; l.v[1] = v.y;
			} else {
				l.v[0] = __builtin_nontemporal_load(p);
			}
		} else {
#pragma unroll
			for (int r = 0; r < R; ++r)
				if (lrow + r < nrows)
					l.v[r] = p[r];
		}
	} else if (c.width == 8) {
		const u32x2 *p = (const u32x2 *)base + lrow;
		if (full && R % 2 == 0) {
#pragma unroll
			for (int q = 0; q < R / 2; ++q) {
				const u32x4 v = __builtin_nontemporal_load((const u32x4 *)p + q);
				l.v[2 * q] = v.x; h.v[2 * q] = v.y; l.v[2 * q + 1] = v.z; h.v[2 * q + 1] = v.w;
			}
		} else {
#pragma unroll
			for (int r = 0; r < R; ++r)
				if (lrow + r < nrows) {
					const u32x2 v = p[r];
					l.v[r] = v.x;
					h.v[r] = v.y;
				}
		}
	} else {
		const int8_t *p = (const int8_t *)base + lrow;
#pragma unroll
		for (int r = 0; r < R; ++r)
			if (lrow + r < nrows)
				l.v[r] = (uint32_t)(int32_t)p[r];
	}
	lds_put<R>(regs + c.slot * TR + lrow, l);
	if (c.width == 8)
		lds_put<R>(regs + c.slot_hi * TR + lrow, h);
}

template <int R, uint32_t TR>
__device__ __forceinline__ void load_phase(const vg_prog &P, const vg_tabdesc *D, uint32_t row0,
	uint32_t lrow, uint32_t nrows, uint32_t phase, uint32_t *regs)
{
	const bool aligned = D->aligned != 0;
	for (uint32_t k = 0; k < P.nload; ++k) {
		const vg_col &c = P.col[k];
		if (c.phase != phase)
			continue;
		const char *base = (const char *)D->col[c.src] + (size_t)row0 * c.width;
		load_col<R, TR>(c, base, lrow, nrows, aligned, regs);
	}
}

/* ----------------------------------------------------------- look-back */
__device__ __forceinline__ uint64_t st_make(uint32_t epoch, uint64_t flag, uint64_t v)
{
	return ((uint64_t)epoch << 40) | (flag << 38) | (v & VG_ST_VALMASK);
}

/* Exclusive prefix of `tile`: publish its aggregate, then walk back over
 * predecessors 64 at a time (one per lane) until one with a published
 * inclusive prefix is found.  Called by one full wave. */
__device__ uint64_t lookback(const vg_launch &L, uint32_t tile, uint32_t agg, uint32_t lane)
{
	uint64_t *st = L.status;
	if (tile == 0) {
		if (lane == 0)
			__hip_atomic_store(&st[0], st_make(L.epoch, VG_ST_PREFIX, agg),
				__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		return 0;
	}
	if (lane == 0)
		__hip_atomic_store(&st[tile], st_make(L.epoch, VG_ST_AGG, agg),
			__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	uint64_t excl = 0;
	int64_t j = (int64_t)tile - 1;
	uint32_t spins = 0;
	while (true) {
		const int64_t idx = j - (int64_t)lane;
		uint64_t flag = VG_ST_PREFIX, val = 0;
		if (idx >= 0) {
			const uint64_t w = __hip_atomic_load(&st[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			flag = (uint32_t)(w >> 40) == L.epoch ? (w >> 38) & 3ull : 0ull;
			val = w & VG_ST_VALMASK;
		}
		const uint64_t pmask = __ballot(flag == VG_ST_PREFIX);
		const uint64_t xmask = __ballot(flag == 0);
		const uint32_t first_p = pmask ? (uint32_t)__builtin_ctzll(pmask) : 64u;
		const uint64_t upto = first_p >= 63 ? ~0ull : ((2ull << first_p) - 1);
		if (xmask & upto) {
			if (++spins > VG_SPIN_LIMIT) {
				if (lane == 0)
					atomicOr(&L.ctrl->error, 1u);
				break;
			}
			__builtin_amdgcn_s_sleep(2);
			continue;
		}
		uint64_t part = lane <= first_p ? val : 0;
#pragma unroll
		for (int off = 32; off > 0; off >>= 1)
			part += __shfl_xor(part, off, 64);
		excl += part;
		if (first_p < 64)
			break;
		j -= 64;
	}
	if (lane == 0)
		__hip_atomic_store(&st[tile], st_make(L.epoch, VG_ST_PREFIX, excl + agg),
			__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	return excl;
}

/* --------------------------------------------------------------- kernel */
/* dynamic LDS: nslot * T words of registers, then T 16-bit row numbers */
template <int R, int NT>
__global__ __launch_bounds__(NT) void vg_select_kernel(const vg_prog P, const vg_launch L)
{
	constexpr uint32_t T = NT * R;
	constexpr int NW = NT / 64;
	extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
	__shared__ uint32_t s_tile;
	__shared__ uint32_t s_wcount[NW];
	__shared__ uint64_t s_base;
	uint32_t *regs = lds;
	uint16_t *srcrow = (uint16_t *)(lds + P.nslot * T);

	const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	if (tid == 0)
		s_tile = atomicAdd(&L.ctrl->tile_counter, 1u);
	__syncthreads();
	const uint32_t tile = uni(s_tile);
	if (tile >= L.num_tiles) {
		/* the counter was not reset by an earlier launch: report, touch nothing */
		if (tid == 0)
			atomicOr(&L.ctrl->error, 2u);
		return;
	}
	const uint32_t tb = uni(L.tile_tab[tile]);
	const vg_tabdesc *D = &L.tabs[tb];
	const uint32_t row0 = (tile - D->first_tile) * T;
	const uint32_t left = D->rows - row0;
	const uint32_t nrows = left < T ? left : T;
	const uint32_t lrow = (wave * 64 + lane) * R;

	uint32_t resume[R];
	bool valid[R];
#pragma unroll
	for (int r = 0; r < R; ++r) {
		resume[r] = 0;
		valid[r] = lrow + r < nrows;
	}

	/* every LDS row a lane touches is its own, so no barrier is needed
	 * between the register-file phases */
	load_phase<R, T>(P, D, row0, lrow, nrows, VG_PHASE_FILTER, regs);
	run_uops<R, T>(P, 0, P.nfilter, regs, lrow, resume, valid);

	if (P.nuop > P.nfilter || P.has_project_loads) {
		bool lane_any = false;
#pragma unroll
		for (int r = 0; r < R; ++r)
			lane_any = lane_any || valid[r];
		if (__ballot(lane_any) != 0) {
			if (lane_any)
				load_phase<R, T>(P, D, row0, lrow, nrows, VG_PHASE_PROJECT, regs);
			run_uops<R, T>(P, P.nfilter, P.nuop, regs, lrow, resume, valid);
		}
	}

	/* ranks inside the wave and the tile */
	uint32_t lane_pref = 0, wave_tot = 0;
#pragma unroll
	for (int r = 0; r < R; ++r) {
		const uint64_t b = __ballot(valid[r]);
		lane_pref += rank_below(b);
		wave_tot += (uint32_t)__builtin_popcountll(b);
	}
	if (lane == 0)
		s_wcount[wave] = wave_tot;
	__syncthreads();
	uint32_t wave_off = 0, tile_tot = 0;
#pragma unroll
	for (int w = 0; w < NW; ++w) {
		const uint32_t c = s_wcount[w];
		wave_off += (uint32_t)w < wave ? c : 0u;
		tile_tot += c;
	}
	{
		uint32_t pos = wave_off + lane_pref;
#pragma unroll
		for (int r = 0; r < R; ++r)
			if (valid[r])
				srcrow[pos++] = (uint16_t)(lrow + r);
	}
	if (wave == 0) {
		const uint64_t base = lookback(L, tile, tile_tot, lane);
		if (lane == 0) {
			s_base = base;
			if (tile == L.num_tiles - 1)
				L.ctrl->total = base + tile_tot;
			if (base + tile_tot > L.capacity)
				atomicOr(&L.ctrl->overflow, 1u);
		}
	}
	__syncthreads();
	const uint64_t base = s_base;
	const uint64_t room = L.capacity > base ? L.capacity - base : 0;
	const uint32_t nwrite = (uint64_t)tile_tot < room ? tile_tot : (uint32_t)room;

	/* gather every result column in row order */
	for (uint32_t j = 0; j < P.nout; ++j) {
		const vg_out &o = P.out[j];
		char *dst = (char *)L.out[j];
		if (o.kind == VG_OUT_SLOT) {
			const uint32_t *lo = regs + o.slot * T, *hi = regs + o.slot_hi * T;
			if (o.width == 8) {
				for (uint32_t i = tid; i < nwrite; i += NT) {
					const uint32_t s = srcrow[i];
					((uint64_t *)dst)[base + i] = ((uint64_t)hi[s] << 32) | lo[s];
				}
			} else if (o.width == 4) {
				for (uint32_t i = tid; i < nwrite; i += NT)
					((uint32_t *)dst)[base + i] = lo[srcrow[i]];
			} else {
				for (uint32_t i = tid; i < nwrite; i += NT)
					((uint8_t *)dst)[base + i] = (uint8_t)lo[srcrow[i]];
			}
		} else if (o.kind == VG_OUT_COLUMN) {
			const vg_col &c = P.col[o.col];
			const char *src = (const char *)D->col[c.src] + (size_t)row0 * c.width;
			if (o.width == 8) {
				for (uint32_t i = tid; i < nwrite; i += NT)
					((uint64_t *)dst)[base + i] = ((const uint64_t *)src)[srcrow[i]];
			} else if (o.width == 4) {
				for (uint32_t i = tid; i < nwrite; i += NT)
					((uint32_t *)dst)[base + i] = ((const uint32_t *)src)[srcrow[i]];
			} else {
				for (uint32_t i = tid; i < nwrite; i += NT)
					((uint8_t *)dst)[base + i] = ((const uint8_t *)src)[srcrow[i]];
			}
		} else {
			if (o.width == 8) {
				for (uint32_t i = tid; i < nwrite; i += NT)
					((uint64_t *)dst)[base + i] = o.imm;
			} else if (o.width == 4) {
				for (uint32_t i = tid; i < nwrite; i += NT)
					((uint32_t *)dst)[base + i] = (uint32_t)o.imm;
			} else {
				for (uint32_t i = tid; i < nwrite; i += NT)
					((uint8_t *)dst)[base + i] = (uint8_t)o.imm;
			}
		}
	}

	__syncthreads();
	if (tid == 0) {
		__threadfence();
		const uint32_t d = atomicAdd(&L.ctrl->done, 1u);
		if (d == L.num_tiles - 1) {
			L.ctrl->tile_counter = 0;
			L.ctrl->done = 0;
		}
	}
}

/* --------------------------------------------------------------- launch */
template <int R>
static int launch_r(const vg_prog *P, const vg_launch *L, hipStream_t stream)
{
	constexpr uint32_t T = 256u * R;
	const size_t lds = (size_t)P->nslot * T * 4 + T * 2;
	hipLaunchKernelGGL((vg_select_kernel<R, 256>), dim3(L->num_tiles), dim3(256), lds, stream, *P, *L);
	return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" size_t vg_kernel_lds_bytes(const vg_prog *P, int rows_per_lane)
{
	const size_t T = 256u * (size_t)rows_per_lane;
	return (size_t)P->nslot * T * 4 + T * 2;
}

#define VG_LDS_LIMIT (64u * 1024u)

/* Rows per lane to use for a program: the requested value, halved until
 * its register file fits the workgroup LDS budget. */
extern "C" int vg_kernel_pick_rpl(const vg_prog *P, int requested)
{
	int r = requested == 1 || requested == 2 || requested == 4 ? requested : 4;
	while (r > 1 && vg_kernel_lds_bytes(P, r) > VG_LDS_LIMIT)
		r /= 2;
	return vg_kernel_lds_bytes(P, r) > VG_LDS_LIMIT ? 0 : r;
}

extern "C" int vg_kernel_launch(const vg_prog *P, const vg_launch *L, int rows_per_lane,
	hipStream_t stream)
{
	if (vg_kernel_lds_bytes(P, rows_per_lane) > VG_LDS_LIMIT)
		return -3;
	switch (rows_per_lane) {
	case 1: return launch_r<1>(P, L, stream);
	case 2: return launch_r<2>(P, L, stream);
	case 4: return launch_r<4>(P, L, stream);
	}
	return -1;
}

extern "C" uint32_t vg_kernel_tile_rows(int rows_per_lane)
{
	return 256u * (uint32_t)rows_per_lane;
}
