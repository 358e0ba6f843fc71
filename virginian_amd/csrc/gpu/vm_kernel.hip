/*
 * gfx950 kernels of the opcode VM (the data-parallel section of a query:
 * reference src/vm/vm_gpu.cu:1152-1225 on the GPU, src/vm/virginia.c:228-891
 * on the CPU).
 *
 * A query is two launches over the same persistent grid.  Rows are cut
 * into tiles of T = NT*R consecutive rows of one tablet; workgroup g owns
 * a contiguous run of tiles.  Inside a tile, lane l of wave w owns rows
 * (w*64+l)*R .. +R-1, so a column load is one contiguous R*width-byte
 * vector per lane and a wave reads 64*R*width contiguous bytes of HBM.
 *
 *   count pass: load the columns the filter micro-ops read, run them,
 *     store one validity byte per lane per tile and one row count per
 *     workgroup;
 *   write pass: a workgroup's output offset is the sum of the counts of
 *     the workgroups before it.  Per tile, each wave gets its own offset
 *     from the tile's validity bytes, loads the columns the result needs
 *     only in lanes holding a selected row, runs the micro-ops computing
 *     result registers, ranks its rows with ballot + mbcnt and writes its
 *     contiguous slice of every result column in row order.
 * Nothing waits on another workgroup, so neither pass has atomics, spins
 * or cross-workgroup ordering assumptions; results come out in the
 * reference's single-core row order.
 *
 * The VM registers live in LDS as [slot][row of the tile]; a lane's R rows
 * of a slot are contiguous, so reading an operand for all of them is one
 * ds_read_b128 (R = 4) at a wave-uniform slot base.  Micro-op fields are
 * read with scalar loads from the kernel arguments, so dispatch is a
 * scalar branch.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "vg_prog.h"

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

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
__device__ __forceinline__ void run_uops(const vg_prog &P, uint32_t kbeg, uint32_t kend, uint64_t mask,
	uint32_t *regs, uint32_t lrow, uint32_t (&resume)[R], bool (&valid)[R])
{
	for (uint32_t k = kbeg; k < kend; ++k) {
		if (!((mask >> k) & 1ull))
			continue;
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
__device__ __forceinline__ void load_cols(const vg_prog &P, const vg_tabdesc *D, uint32_t row0,
	uint32_t lrow, uint32_t nrows, bool write_pass, uint32_t *regs)
{
	const bool aligned = D->aligned != 0;
	for (uint32_t k = 0; k < P.nload; ++k) {
		const vg_col &c = P.col[k];
		if (write_pass ? !c.wload : c.phase != VG_PHASE_FILTER)
			continue;
		const char *base = (const char *)D->col[c.src] + (size_t)row0 * c.width;
		load_col<R, TR>(c, base, lrow, nrows, aligned, regs);
	}
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		x += __shfl_xor(x, off, 64);
	return x;
}

/* ----------------------------------------------------------- count pass */
/* Workgroup g evaluates the filter micro-ops over its tiles, stores one
 * validity byte per lane per tile (bit r = row lrow+r) and the number of
 * selected rows of all its tiles.  LDS: nslot * T words of registers. */
template <int R, int NT>
__global__ __launch_bounds__(NT) void vg_count_kernel(const vg_prog P, const vg_launch L)
{
	constexpr uint32_t T = NT * R;
	constexpr int NW = NT / 64;
	extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
	__shared__ uint32_t s_wcount[NW];
	uint32_t *regs = lds;
	const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const uint32_t lrow = (wave * 64 + lane) * R;
	const uint32_t g = blockIdx.x;
	const uint32_t t0 = g * L.tiles_per_wg;
	const uint32_t t1 = min(t0 + L.tiles_per_wg, L.num_tiles);
	uint32_t count = 0;
	for (uint32_t tile = t0; tile < t1; ++tile) {
		const uint32_t tb = uni(L.tile_tab[tile]);
		const vg_tabdesc *D = &L.tabs[tb];
		const uint32_t row0 = (tile - D->first_tile) * T;
		const uint32_t left = D->rows - row0;
		const uint32_t nrows = left < T ? left : T;
		uint32_t resume[R];
		bool valid[R];
#pragma unroll
		for (int r = 0; r < R; ++r) {
			resume[r] = 0;
			valid[r] = lrow + r < nrows;
		}
		/* every LDS row a lane touches is its own: no barriers needed */
		load_cols<R, T>(P, D, row0, lrow, nrows, false, regs);
		run_uops<R, T>(P, 0, P.nfilter, ~0ull, regs, lrow, resume, valid);
		uint32_t bits = 0;
#pragma unroll
		for (int r = 0; r < R; ++r)
			bits |= valid[r] ? (1u << r) : 0u;
		L.vbits[(size_t)tile * NT + tid] = (uint8_t)bits;
		count += (uint32_t)__builtin_popcount(bits);
	}
	count = wave_sum(count);
	if (lane == 0)
		s_wcount[wave] = count;
	__syncthreads();
	if (tid == 0) {
		uint32_t s = 0;
#pragma unroll
		for (int w = 0; w < NW; ++w)
			s += s_wcount[w];
		L.wgcount[g] = s;
	}
}

/* ----------------------------------------------------------- write pass */
/* Workgroup g starts at the sum of the counts of workgroups < g.  For each
 * of its tiles the four waves compute their offsets from the tile's
 * validity bytes; a wave with selected rows loads what the result needs
 * (only in lanes holding a selected row), runs the micro-ops that compute
 * result registers, ranks its rows with ballot + mbcnt, and writes its
 * contiguous slice of every result column.  Waves never wait for each
 * other.  LDS: nslot * T words of registers, then R*64 row numbers per
 * wave. */
template <int R, int NT>
__global__ __launch_bounds__(NT) void vg_write_kernel(const vg_prog P, const vg_launch L)
{
	constexpr uint32_t T = NT * R;
	constexpr int NW = NT / 64;
	static_assert(NT / 4 == 64, "one 4-byte word of validity bytes per lane covers the tile");
	extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
	uint32_t *regs = lds;
	const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	uint16_t *srcrow = (uint16_t *)(lds + P.nslot * T) + wave * (64 * R);
	const uint32_t lrow = (wave * 64 + lane) * R;
	const uint32_t g = blockIdx.x;
	const uint32_t t0 = g * L.tiles_per_wg;
	const uint32_t t1 = min(t0 + L.tiles_per_wg, L.num_tiles);

	/* offset of this workgroup: sum of the counts before it */
	uint64_t base = 0;
	for (uint32_t i = lane; i < g; i += 64)
		base += L.wgcount[i];
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		base += __shfl_xor(base, off, 64);

	for (uint32_t tile = t0; tile < t1; ++tile) {
		/* validity bytes of the whole tile: lane l holds threads 4l..4l+3 */
		const uint32_t word = ((const uint32_t *)(L.vbits + (size_t)tile * NT))[lane];
		const uint32_t c = (uint32_t)__builtin_popcount(word);
		const uint32_t tile_tot = wave_sum(c);
		if (tile_tot == 0)
			continue;
		const uint32_t wave_off = wave_sum(lane < 16 * wave ? c : 0u);
		const uint32_t wave_tot = wave_sum(lane >= 16 * wave && lane < 16 * wave + 16 ? c : 0u);
		if (wave_tot != 0) {
			const uint32_t mybits = (__shfl(word, (int)(16 * wave + (lane >> 2)), 64) >> ((lane & 3) * 8)) & 0xFFu;
			const uint32_t tb = uni(L.tile_tab[tile]);
			const vg_tabdesc *D = &L.tabs[tb];
			const uint32_t row0 = (tile - D->first_tile) * T;
			const uint32_t left = D->rows - row0;
			const uint32_t nrows = left < T ? left : T;
			uint32_t resume[R];
			bool valid[R];
#pragma unroll
			for (int r = 0; r < R; ++r) {
				resume[r] = 0;
				valid[r] = (mybits >> r) & 1u;
			}
			if (mybits)
				load_cols<R, T>(P, D, row0, lrow, nrows, true, regs);
			if (P.wfull) {
				/* rerun everything: resume points and validity come back
				 * exactly as in the count pass */
				run_uops<R, T>(P, 0, P.nuop, ~0ull, regs, lrow, resume, valid);
			} else if (P.wneed) {
				run_uops<R, T>(P, 0, P.nuop, P.wneed, regs, lrow, resume, valid);
			}
			/* ranks inside the wave, row numbers in rank order */
			uint32_t pos = 0;
#pragma unroll
			for (int r = 0; r < R; ++r) {
				const uint64_t b = __ballot((mybits >> r) & 1u);
				pos += rank_below(b);
			}
#pragma unroll
			for (int r = 0; r < R; ++r)
				if ((mybits >> r) & 1u)
					srcrow[pos++] = (uint16_t)(lrow + r);
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

			const uint64_t at = base + wave_off;
			const uint64_t room = L.capacity > at ? L.capacity - at : 0;
			const uint32_t n = (uint64_t)wave_tot < room ? wave_tot : (uint32_t)room;
			if (n < wave_tot && lane == 0)
				atomicOr(&L.ctrl->overflow, 1u);
			for (uint32_t j = 0; j < P.nout; ++j) {
				const vg_out &o = P.out[j];
				char *dst = (char *)L.out[j];
				if (o.kind == VG_OUT_SLOT) {
					const uint32_t *lo = regs + o.slot * T, *hi = regs + o.slot_hi * T;
					if (o.width == 8) {
						for (uint32_t i = lane; i < n; i += 64) {
							const uint32_t s = srcrow[i];
							((uint64_t *)dst)[at + i] = ((uint64_t)hi[s] << 32) | lo[s];
						}
					} else if (o.width == 4) {
						for (uint32_t i = lane; i < n; i += 64)
							((uint32_t *)dst)[at + i] = lo[srcrow[i]];
					} else {
						for (uint32_t i = lane; i < n; i += 64)
							((uint8_t *)dst)[at + i] = (uint8_t)lo[srcrow[i]];
					}
				} else if (o.kind == VG_OUT_COLUMN) {
					const vg_col &cc = P.col[o.col];
					const char *src = (const char *)D->col[cc.src] + (size_t)row0 * cc.width;
					if (o.width == 8) {
						for (uint32_t i = lane; i < n; i += 64)
							((uint64_t *)dst)[at + i] = ((const uint64_t *)src)[srcrow[i]];
					} else if (o.width == 4) {
						for (uint32_t i = lane; i < n; i += 64)
							((uint32_t *)dst)[at + i] = ((const uint32_t *)src)[srcrow[i]];
					} else {
						for (uint32_t i = lane; i < n; i += 64)
							((uint8_t *)dst)[at + i] = ((const uint8_t *)src)[srcrow[i]];
					}
				} else {
					if (o.width == 8) {
						for (uint32_t i = lane; i < n; i += 64)
							((uint64_t *)dst)[at + i] = o.imm;
					} else if (o.width == 4) {
						for (uint32_t i = lane; i < n; i += 64)
							((uint32_t *)dst)[at + i] = (uint32_t)o.imm;
					} else {
						for (uint32_t i = lane; i < n; i += 64)
							((uint8_t *)dst)[at + i] = (uint8_t)o.imm;
					}
				}
			}
			/* srcrow and the register rows are rewritten by the next tile */
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		}
		base += tile_tot;
	}
	if (g == gridDim.x - 1 && tid == 0)
		L.ctrl->total = base;
}

/* --------------------------------------------------------------- launch */
template <int R>
static size_t lds_bytes(const vg_prog *P)
{
	constexpr size_t T = 256u * R;
	return (size_t)P->nslot * T * 4 + 4 * (64 * R) * 2;
}

extern "C" size_t vg_kernel_lds_bytes(const vg_prog *P, int rows_per_lane)
{
	switch (rows_per_lane) {
	case 1: return lds_bytes<1>(P);
	case 2: return lds_bytes<2>(P);
	default: return lds_bytes<4>(P);
	}
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

/* Workgroups per pass: enough to fill every CU at the occupancy the LDS
 * footprint allows, never more than there are tiles. */
template <int R>
static uint32_t grid_for(const vg_prog *P, uint32_t num_tiles)
{
	static int cus = 0;
	if (cus == 0) {
		int dev = 0;
		hipDeviceProp_t prop;
		if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
			cus = 256;
		else
			cus = prop.multiProcessorCount;
	}
	int per_cu = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, vg_write_kernel<R, 256>, 256,
			lds_bytes<R>(P)) != hipSuccess || per_cu < 1)
		per_cu = 1;
	uint32_t g = (uint32_t)cus * (uint32_t)per_cu;
	return g < num_tiles ? g : num_tiles;
}

template <int R>
static int launch_r(const vg_prog *P, vg_launch *L, hipStream_t stream)
{
	const uint32_t g = grid_for<R>(P, L->num_tiles);
	L->tiles_per_wg = (L->num_tiles + g - 1) / g;
	const uint32_t grid = (L->num_tiles + L->tiles_per_wg - 1) / L->tiles_per_wg;
	const size_t lds = lds_bytes<R>(P);
	hipLaunchKernelGGL((vg_count_kernel<R, 256>), dim3(grid), dim3(256), lds, stream, *P, *L);
	if (hipGetLastError() != hipSuccess)
		return -2;
	hipLaunchKernelGGL((vg_write_kernel<R, 256>), dim3(grid), dim3(256), lds, stream, *P, *L);
	return hipGetLastError() == hipSuccess ? 0 : -2;
}

/* Number of workgroups a launch of this program over num_tiles tiles uses
 * (the size of the per-workgroup count array). */
extern "C" uint32_t vg_kernel_grid(const vg_prog *P, int rows_per_lane, uint32_t num_tiles)
{
	if (num_tiles == 0)
		return 0;
	uint32_t g;
	switch (rows_per_lane) {
	case 1: g = grid_for<1>(P, num_tiles); break;
	case 2: g = grid_for<2>(P, num_tiles); break;
	default: g = grid_for<4>(P, num_tiles); break;
	}
	const uint32_t per = (num_tiles + g - 1) / g;
	return (num_tiles + per - 1) / per;
}

extern "C" int vg_kernel_launch(const vg_prog *P, vg_launch *L, int rows_per_lane,
	hipStream_t stream)
{
	if (L->num_tiles == 0)
		return 0;
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
