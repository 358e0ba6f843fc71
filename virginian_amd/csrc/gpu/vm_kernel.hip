/*
 * gfx950 kernels of the opcode VM (the data-parallel section of a query:
 * reference src/vm/vm_gpu.cu:1152-1225 on the GPU, src/vm/virginia.c:228-891
 * on the CPU).
 *
 * A query is two launches over the same persistent grid.  Rows are cut
 * into tiles of 1024 consecutive rows of one tablet; workgroup g owns a
 * contiguous run of tiles.  Rows are handled in strips of 4 consecutive
 * rows per lane, so a wave's strip is 256 rows: one 1 KB contiguous
 * access in HBM and in LDS.
 *
 *   count pass: each wave takes a whole tile (4 strips per lane), copies
 *     the columns the filter micro-ops read from HBM straight into its LDS
 *     register file, runs those micro-ops, and stores one validity byte per
 *     write-pass lane, the selected-row count of each write-pass wave, and
 *     one count per workgroup;
 *   write pass: a workgroup's output offset is the sum of the counts of
 *     the workgroups before it, and each tile's wave offsets follow from
 *     the per-wave counts.  A wave with selected rows loads what the result
 *     needs (only in lanes holding a selected row), runs the micro-ops
 *     computing result registers, ranks its rows with ballot + mbcnt and
 *     writes its contiguous slice of every result column in row order.
 * Nothing waits on another workgroup, so neither pass has atomics, spins
 * or cross-workgroup ordering assumptions; results come out in the
 * reference's single-core row order.
 *
 * The VM registers live in LDS as [slot][row of the tile]; reading an
 * operand for a strip is one ds_read_b128 at a wave-uniform slot base.
 * Micro-op fields are read with scalar loads from the kernel arguments,
 * so dispatch is a scalar branch.
 */
#include <hip/hip_runtime.h>
#include <unordered_map>
#include <stdint.h>
#include "vg_prog.h"

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t uni(uint32_t x)
{
	return __builtin_amdgcn_readfirstlane(x);
}

/* Read memory that no kernel of the launch writes (tile table, tablet
 * descriptors) through the constant address space: with a wave-uniform
 * address this is a scalar load, which neither occupies VGPRs nor waits
 * behind the vector loads already in flight. */
template <typename T>
__device__ __forceinline__ T ldc(const T *p)
{
	return *(const __attribute__((address_space(4))) T *)p;
}

/* The same for a small struct of byte fields (micro-ops, column and
 * output descriptors): fetched as whole dwords, so the fields are
 * extracted from SGPRs instead of being read by vector byte loads. */
template <typename T>
__device__ __forceinline__ T ldc_struct(const T *p)
{
	static_assert(sizeof(T) % 4 == 0, "dword-sized descriptors only");
	union {
		T t;
		uint32_t w[sizeof(T) / 4];
	} u;
	const __attribute__((address_space(4))) uint32_t *s =
		(const __attribute__((address_space(4))) uint32_t *)p;
#pragma unroll
	for (unsigned i = 0; i < sizeof(T) / 4; ++i)
		u.w[i] = s[i];
	return u.t;
}

/* global (not flat) view of column data, so vector loads count only
 * against vmcnt */
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gptr(const void *p)
{
	return (const __attribute__((address_space(1))) T *)p;
}

/* sum of x over the wave, in every lane */
__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		x += __shfl_xor(x, off, 64);
	return x;
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

/* ------------------------------------------------------ strips of rows */
/* A lane owns NS strips of 4 consecutive rows.  Strip q of a lane starts
 * at row wbase + q*stride + 4*lane of the tile: wbase and stride are
 * wave-uniform, so a wave's strip q is 256 consecutive rows, which is one
 * 16-byte-per-lane vector access both in HBM and in the LDS register
 * file ([slot][row of the tile]). */
template <int NS>
struct strips {
	uint32_t wbase, stride, lane;
	__device__ __forceinline__ uint32_t row(int q) const { return wbase + (uint32_t)q * stride + 4 * lane; }
	__device__ __forceinline__ uint32_t ubase(int q) const { return wbase + (uint32_t)q * stride; }
};

/* ------------------------------------------------------------ row state */
/* The rows of a lane as bits: bit 4q+j is row j of strip q.  `active`
 * rows execute the micro-op at hand (the reference's row_pc == pc).  A row
 * that takes a forward branch keeping its validity waits in wait[slot]
 * until the branch target, where it is active again.  A row that loses
 * validity can never reach the result (validity only ever drops,
 * virginia.c:59-62), so it leaves `active` for good and what it would
 * compute afterwards is irrelevant.  With rows as bit masks every
 * micro-op is straight-line vector code: no per-row branches, hence no
 * exec-mask bookkeeping on the scalar unit. */
struct rowstate {
	uint32_t valid, active;
	uint32_t wait[VG_NWAIT];
};

__device__ __forceinline__ void rows_init(rowstate &RS, uint32_t valid)
{
	RS.valid = RS.active = valid;
#pragma unroll
	for (int j = 0; j < VG_NWAIT; ++j)
		RS.wait[j] = 0;
}

/* ------------------------------------------------------------- micro-ops */
/* condition bits of a compare / logic micro-op over the lane's rows */
template <typename T, int NS, uint32_t TR, bool USE_B, typename F>
__device__ __forceinline__ uint32_t cond_strips(const vg_uop &u, const uint32_t *regs,
	const strips<NS> &S, F f)
{
	uint32_t c = 0;
#pragma unroll
	for (int q = 0; q < NS; ++q) {
		T a[4], b[4];
		operand<T, 4, TR>(regs, u.flags & VG_F_AIMM, u.ia, u.a, u.a_hi, S.row(q), a);
		if (USE_B)
			operand<T, 4, TR>(regs, u.flags & VG_F_BIMM, u.ib, u.b, u.b_hi, S.row(q), b);
#pragma unroll
		for (int r = 0; r < 4; ++r)
			c |= f(a[r], USE_B ? b[r] : a[r]) ? 1u << (4 * q + r) : 0u;
	}
	return c;
}

/* write 4 typed values of strip row r0 to a slot; a masked write keeps
 * the old value of every row that is not active (bits `act`) */
template <typename T, uint32_t TR>
__device__ __forceinline__ void result_put(uint32_t *regs, uint32_t lo, uint32_t hi, uint32_t r0,
	const T (&x)[4], bool masked, uint32_t act)
{
	rv<4> l, h;
	if (masked) {
		l = lds_get<4>(regs + lo * TR + r0);
		if constexpr (vt<T>::wide)
			h = lds_get<4>(regs + hi * TR + r0);
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			const bool on = (act >> r) & 1u;
			l.v[r] = on ? vt<T>::lo(x[r]) : l.v[r];
			if constexpr (vt<T>::wide)
				h.v[r] = on ? vt<T>::hi(x[r]) : h.v[r];
		}
	} else {
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			l.v[r] = vt<T>::lo(x[r]);
			if constexpr (vt<T>::wide)
				h.v[r] = vt<T>::hi(x[r]);
		}
	}
	lds_put<4>(regs + lo * TR + r0, l);
	if constexpr (vt<T>::wide)
		lds_put<4>(regs + hi * TR + r0, h);
}

template <typename T, int NS, uint32_t TR, typename F>
__device__ __forceinline__ void math_strips(const vg_uop &u, uint32_t *regs, const strips<NS> &S,
	uint32_t active, F f)
{
	const bool masked = u.flags & VG_F_MASKED;
#pragma unroll
	for (int q = 0; q < NS; ++q) {
		T a[4], b[4], x[4];
		operand<T, 4, TR>(regs, u.flags & VG_F_AIMM, u.ia, u.a, u.a_hi, S.row(q), a);
		operand<T, 4, TR>(regs, u.flags & VG_F_BIMM, u.ib, u.b, u.b_hi, S.row(q), b);
#pragma unroll
		for (int r = 0; r < 4; ++r)
			x[r] = f(a[r], b[r]);
		result_put<T, TR>(regs, u.dst, u.dst_hi, S.row(q), x, masked, active >> (4 * q));
	}
}

template <typename T, int NS, uint32_t TR>
__device__ __forceinline__ void mov_strips(const vg_uop &u, uint32_t *regs, const strips<NS> &S,
	uint32_t active)
{
	const bool masked = u.flags & VG_F_MASKED;
#pragma unroll
	for (int q = 0; q < NS; ++q) {
		T a[4];
		operand<T, 4, TR>(regs, u.flags & VG_F_AIMM, u.ia, u.a, u.a_hi, S.row(q), a);
		result_put<T, TR>(regs, u.dst, u.dst_hi, S.row(q), a, masked, active >> (4 * q));
	}
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

/* Cast converts the register in place for the rows at this op
 * (virginia.c:682-862 tests row_pc <= pc, which only active rows meet) */
template <typename Sx, typename Dx, int NS, uint32_t TR>
__device__ __forceinline__ void cast_strips(const vg_uop &u, uint32_t *regs, const strips<NS> &S,
	uint32_t active)
{
#pragma unroll
	for (int q = 0; q < NS; ++q) {
		Sx a[4];
		Dx x[4];
		operand<Sx, 4, TR>(regs, false, 0, u.a, u.a_hi, S.row(q), a);
#pragma unroll
		for (int r = 0; r < 4; ++r)
			x[r] = convert<Sx, Dx>(a[r]);
		result_put<Dx, TR>(regs, u.dst, u.dst_hi, S.row(q), x, true, active >> (4 * q));
	}
}

template <typename Sx, int NS, uint32_t TR>
__device__ __forceinline__ void cast_from(const vg_uop &u, uint32_t *regs, const strips<NS> &S,
	uint32_t active)
{
	switch (u.b) {
	case VG_T_INT: cast_strips<Sx, int32_t, NS, TR>(u, regs, S, active); break;
	case VG_T_INT64: cast_strips<Sx, int64_t, NS, TR>(u, regs, S, active); break;
	case VG_T_FLOAT: cast_strips<Sx, float, NS, TR>(u, regs, S, active); break;
	case VG_T_DOUBLE: cast_strips<Sx, double, NS, TR>(u, regs, S, active); break;
	case VG_T_CHAR: cast_strips<Sx, chr, NS, TR>(u, regs, S, active); break;
	}
}

/* One case per (kind, type) of a micro-op, so that dispatch is a single
 * scalar compare-and-branch search on kind * 8 + type rather than nested
 * switches on the kind and the type (about 5 % off the write pass).
 * Compares set c and leave the switch for the branch; the others go on
 * with the next micro-op. */
#define VG_FLAT_CASES(TY, T) \
	case VG_K_LT * 8 + TY: c = cond_strips<T, NS, TR, true>(u, regs, S, [](T x, T y) { return x < y; }); break; \
	case VG_K_LE * 8 + TY: c = cond_strips<T, NS, TR, true>(u, regs, S, [](T x, T y) { return x <= y; }); break; \
	case VG_K_GT * 8 + TY: c = cond_strips<T, NS, TR, true>(u, regs, S, [](T x, T y) { return x > y; }); break; \
	case VG_K_GE * 8 + TY: c = cond_strips<T, NS, TR, true>(u, regs, S, [](T x, T y) { return x >= y; }); break; \
	case VG_K_EQ * 8 + TY: c = cond_strips<T, NS, TR, true>(u, regs, S, [](T x, T y) { return x == y; }); break; \
	case VG_K_NE * 8 + TY: c = cond_strips<T, NS, TR, true>(u, regs, S, [](T x, T y) { return x != y; }); break; \
	case VG_K_AND * 8 + TY: c = cond_strips<T, NS, TR, true>(u, regs, S, [](T x, T y) { return nz(x) && nz(y); }); break; \
	case VG_K_OR * 8 + TY: c = cond_strips<T, NS, TR, true>(u, regs, S, [](T x, T y) { return nz(x) || nz(y); }); break; \
	case VG_K_NOT * 8 + TY: c = cond_strips<T, NS, TR, false>(u, regs, S, [](T x, T) { return !nz(x); }); break; \
	case VG_K_ADD * 8 + TY: math_strips<T, NS, TR>(u, regs, S, RS.active, [](T x, T y) { return vt<T>::add(x, y); }); continue; \
	case VG_K_SUB * 8 + TY: math_strips<T, NS, TR>(u, regs, S, RS.active, [](T x, T y) { return vt<T>::sub(x, y); }); continue; \
	case VG_K_MUL * 8 + TY: math_strips<T, NS, TR>(u, regs, S, RS.active, [](T x, T y) { return vt<T>::mul(x, y); }); continue; \
	case VG_K_DIV * 8 + TY: math_strips<T, NS, TR>(u, regs, S, RS.active, [](T x, T y) { return vt<T>::div(x, y); }); continue; \
	case VG_K_MOV * 8 + TY: mov_strips<T, NS, TR>(u, regs, S, RS.active); continue; \
	case VG_K_CAST * 8 + TY: cast_from<T, NS, TR>(u, regs, S, RS.active); continue;

/* Micro-ops [kbeg, kend) whose bit is set in `mask`, over the lane's
 * rows.  Dispatch is scalar: every field of a micro-op is wave-uniform. */
template <int NS, uint32_t TR>
__device__ __forceinline__ void run_uops(const vg_prog &P, uint32_t kbeg, uint32_t kend, uint64_t mask,
	uint32_t *regs, const strips<NS> &S, rowstate &RS)
{
	for (uint32_t k = kbeg; k < kend; ++k) {
		if (!((mask >> k) & 1ull))
			continue;
		const vg_uop u = ldc_struct(&P.uop[k]);
		if (u.join) {
#pragma unroll
			for (int j = 0; j < VG_NWAIT; ++j)
				if ((u.join >> j) & 1u) {
					RS.active |= RS.wait[j];
					RS.wait[j] = 0;
				}
		}
		if (u.kind == VG_K_INVALID) {
			RS.valid &= ~RS.active;
			RS.active = 0;
			continue;
		}
		uint32_t c = 0;
		switch ((uint32_t)u.kind * 8u + u.type) {
			VG_FLAT_CASES(VG_T_INT, int32_t)
			VG_FLAT_CASES(VG_T_INT64, int64_t)
			VG_FLAT_CASES(VG_T_FLOAT, float)
			VG_FLAT_CASES(VG_T_DOUBLE, double)
			VG_FLAT_CASES(VG_T_CHAR, chr)
		default:
			continue;
		}
		/* a compare: taken rows jump to the target, kept rows wait for
		 * it, the others lose validity (virginia.c:59-62) */
		const uint32_t taken = c & RS.active;
		RS.active &= ~taken;
		if (u.flags & VG_F_KEEP) {
#pragma unroll
			for (int j = 0; j < VG_NWAIT; ++j)
				if (u.wslot == j)
					RS.wait[j] |= taken;
		} else {
			RS.valid &= ~taken;
		}
	}
}

/* ---------------------------------------------------------- column loads */
typedef __attribute__((address_space(3))) void lds_void;

/* Strips of column c into its LDS slot(s).  A full strip of a 16-byte
 * aligned 4-byte column is copied from HBM straight into the register
 * file (global_load_lds: no VGPRs, and every column of a tile is in
 * flight at once; the compiler waits for them before the first LDS read).
 * 8-byte and 1-byte columns, tails and unaligned tablets go through VGPRs
 * (8-byte values are split into a low-word and a high-word slot). */
template <int NS, uint32_t TR>
__device__ __forceinline__ void load_col(const vg_col &c, const char *base, const strips<NS> &S,
	uint32_t nrows, bool aligned, uint32_t *regs)
{
#pragma unroll
	for (int q = 0; q < NS; ++q) {
		const uint32_t r0 = S.row(q);
		if (c.width == 4 && aligned && r0 + 4 <= nrows) {
			__builtin_amdgcn_global_load_lds((const void *)((const uint32_t *)base + r0),
				(lds_void *)(regs + c.slot * TR + S.ubase(q)), 16, 0, 0);
			continue;
		}
		rv<4> l, h;
#pragma unroll
		for (int r = 0; r < 4; ++r)
			l.v[r] = h.v[r] = 0;
		if (c.width == 4) {
#pragma unroll
			for (int r = 0; r < 4; ++r)
				if (r0 + r < nrows)
					l.v[r] = gptr<uint32_t>(base)[r0 + r];
		} else if (c.width == 8) {
			const __attribute__((address_space(1))) u32x2 *p = gptr<u32x2>(base) + r0;
			if (aligned && r0 + 4 <= nrows) {
#pragma unroll
				for (int j = 0; j < 2; ++j) {
					const u32x4 v = __builtin_nontemporal_load((const __attribute__((address_space(1))) u32x4 *)p + j);
					l.v[2 * j] = v.x; h.v[2 * j] = v.y; l.v[2 * j + 1] = v.z; h.v[2 * j + 1] = v.w;
				}
			} else {
#pragma unroll
				for (int r = 0; r < 4; ++r)
					if (r0 + r < nrows) {
						const u32x2 v = p[r];
						l.v[r] = v.x;
						h.v[r] = v.y;
					}
			}
		} else {
			const __attribute__((address_space(1))) int8_t *p = gptr<int8_t>(base) + r0;
#pragma unroll
			for (int r = 0; r < 4; ++r)
				if (r0 + r < nrows)
					l.v[r] = (uint32_t)(int32_t)p[r];
		}
		lds_put<4>(regs + c.slot * TR + r0, l);
		if (c.width == 8)
			lds_put<4>(regs + c.slot_hi * TR + r0, h);
	}
}

/* rows per tile */
#define VG_T 1024u

/* Wait for every vector memory access of this wave (loads into VGPRs and
 * into LDS alike).  Through the intrinsic the compiler's own wait
 * bookkeeping sees that nothing is outstanding afterwards; the empty asm
 * keeps memory accesses from being moved across it. */
__device__ __forceinline__ void wait_vmem()
{
	__builtin_amdgcn_s_waitcnt(0x0F70);   /* vmcnt(0); expcnt, lgkmcnt unconstrained */
	asm volatile("" ::: "memory");
}

/* ------------------------------------------------------- tablet cursor */
/* Where a wave's current tile lives, kept across tiles so that the chain
 * tile table -> tablet descriptor -> column address is walked once per
 * tablet instead of once per tile.  Lane k of cd_lo/cd_hi holds column
 * descriptor P.col[k] (8 bytes) and lane k of cb_lo/cb_hi the address of
 * that column's row 0 in the current tablet; readlane with a wave-uniform
 * k brings either into SGPRs without a memory access. */
struct cursor {
	uint32_t first, end, rows, aligned;   /* tiles [first, end) of the tablet */
	uint32_t cd_lo, cd_hi, cb_lo, cb_hi;  /* per-lane: column k */
};

__device__ __forceinline__ void cursor_init(cursor &C, const vg_prog &P, uint32_t lane)
{
	C.first = C.end = C.rows = C.aligned = 0;
	C.cd_lo = C.cd_hi = C.cb_lo = C.cb_hi = 0;
	if (lane < ldc(&P.ncol)) {
		const uint32_t *w = (const uint32_t *)&P.col[lane];
		C.cd_lo = w[0];
		C.cd_hi = w[1];
	}
}

/* make `tile` the cursor's tile (tiles only move forward) */
__device__ __forceinline__ void cursor_seek(cursor &C, const vg_prog &P, const vg_launch &L,
	uint32_t tile, uint32_t lane)
{
	if (tile < C.end)
		return;
	const uint32_t tb = ldc(ldc(&L.tile_tab) + tile);
	const vg_tabdesc *D = ldc(&L.tabs) + tb;
	C.first = ldc(&D->first_tile);
	C.rows = ldc(&D->rows);
	C.end = C.first + (C.rows + VG_T - 1) / VG_T;
	C.aligned = ldc(&D->aligned);
	uint64_t b = 0;
	if (lane < ldc(&P.ncol))
		b = gptr<uint64_t>(&D->col[0])[C.cd_lo & 0xFFu];   /* vg_col.src is byte 0 */
	C.cb_lo = (uint32_t)b;
	C.cb_hi = (uint32_t)(b >> 32);
	wait_vmem();
}

__device__ __forceinline__ vg_col cursor_col(const cursor &C, uint32_t k)
{
	union {
		vg_col c;
		uint32_t w[2];
	} u;
	u.w[0] = __builtin_amdgcn_readlane(C.cd_lo, k);
	u.w[1] = __builtin_amdgcn_readlane(C.cd_hi, k);
	return u.c;
}

__device__ __forceinline__ const char *cursor_base(const cursor &C, uint32_t k)
{
	/* readlane yields an int: widen both halves unsigned, or a low word
	 * with bit 31 set would be sign-extended over the high word */
	const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(C.cb_lo, k);
	const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(C.cb_hi, k);
	return (const char *)(((uint64_t)hi << 32) | (uint64_t)lo);
}

/* which columns load_cols loads */
enum { VG_LOAD_FILTER = 0, VG_LOAD_WRITE = 1, VG_LOAD_PROJECT = 2 };

template <int NS, uint32_t TR>
__device__ __forceinline__ void load_cols(const vg_prog &P, const cursor &C, uint32_t row0,
	const strips<NS> &S, uint32_t nrows, int mode, uint32_t *regs, bool wait = true)
{
	const bool aligned = C.aligned != 0;
	for (uint32_t k = 0; k < ldc(&P.nload); ++k) {
		const vg_col c = cursor_col(C, k);
		if (mode == VG_LOAD_WRITE ? !c.wload : (c.phase == VG_PHASE_FILTER) != (mode == VG_LOAD_FILTER))
			continue;
		const char *base = cursor_base(C, k) + (size_t)row0 * c.width;
		load_col<NS, TR>(c, base, S, nrows, aligned, regs);
	}
	/* The compiler does not count global_load_lds as a write of the LDS
	 * it later reads, so it inserts no wait before those reads: wait here
	 * unless the caller waits later. */
	if (wait)
		wait_vmem();
}

/* ----------------------------------------------------------- count pass */
/* Validity leaves the count pass as one byte per write-pass lane: byte
 * tile*256 + 64q + l holds rows q*256 + 4l .. +3 of the tile (bit j = row
 * 4l + j), exactly the rows write-pass lane l of wave q owns;
 * tcount[tile*4 + q] counts the selected rows of strip q and wgcount[g]
 * those of workgroup g, which covers tiles [g*tiles_per_wg,
 * (g+1)*tiles_per_wg).
 *   NS = 4 (a filter needing at most VG_COUNT_WIDE_SLOTS register slots):
 *     each wave takes a whole tile, lane l owning rows q*256 + 4l .. +3 of
 *     strip q < 4, so four 1 KB loads per column are in flight per wave;
 *     the waves take every fourth tile; a register file of
 *     nslot_count * T words per wave.
 *   NS = 1 (more slots): the 4 waves share each tile, wave q taking strip
 *     q as in the write pass; one register file of nslot_count * T words
 *     (a wave only touches its own rows of it). */
/* Count tiles [t0, t1) of a launch with the 4 waves of the workgroup;
 * returns the rows selected in them (in every thread). */
template <int NS>
__device__ __forceinline__ uint32_t count_tiles(const vg_prog &P, const vg_launch &L, uint32_t t0,
	uint32_t t1, uint32_t *lds)
{
	constexpr uint32_t T = VG_T;
	__shared__ uint32_t s_wcount[4];
	const uint32_t tid = threadIdx.x, lane = tid & 63, wave = uni(tid >> 6);
	uint32_t *regs = NS == 4 ? lds + wave * ldc(&P.nslot_count) * T : lds;
	const strips<NS> S{NS == 4 ? 0u : wave * 256u, 256u, lane};
	uint32_t count = 0;
	cursor C;
	cursor_init(C, P, lane);
	for (uint32_t tile = t0 + (NS == 4 ? wave : 0u); tile < t1; tile += (NS == 4 ? 4u : 1u)) {
		cursor_seek(C, P, L, tile, lane);
		const uint32_t row0 = (tile - C.first) * T;
		const uint32_t left = C.rows - row0;
		const uint32_t nrows = left < T ? left : T;
		uint32_t live = (1u << (4 * NS)) - 1u;
		if (nrows < T) {
			live = 0;
#pragma unroll
			for (int q = 0; q < NS; ++q)
#pragma unroll
				for (int j = 0; j < 4; ++j)
					live |= S.row(q) + j < nrows ? 1u << (4 * q + j) : 0u;
		}
		rowstate RS;
		rows_init(RS, live);
		load_cols<NS, T>(P, C, row0, S, nrows, VG_LOAD_FILTER, regs);
		run_uops<NS, T>(P, 0, ldc(&P.nfilter), ~0ull, regs, S, RS);
		const uint32_t v = RS.valid;
		if constexpr (NS == 4) {
#pragma unroll
			for (int q = 0; q < 4; ++q)
				ldc(&L.vbits)[(size_t)tile * 256 + 64 * q + lane] = (uint8_t)((v >> (4 * q)) & 0xFu);
			/* per-strip counts summed over the wave, two 16-bit counts a word */
			const uint32_t n01 = wave_sum((uint32_t)__builtin_popcount(v & 0xFu) |
				((uint32_t)__builtin_popcount(v & 0xF0u) << 16));
			const uint32_t n23 = wave_sum((uint32_t)__builtin_popcount(v & 0xF00u) |
				((uint32_t)__builtin_popcount(v & 0xF000u) << 16));
			if (lane < 4)
				ldc(&L.tcount)[(size_t)tile * 4 + lane] = ((lane < 2 ? n01 : n23) >> (16 * (lane & 1))) & 0xFFFFu;
			count += (n01 & 0xFFFFu) + (n01 >> 16) + (n23 & 0xFFFFu) + (n23 >> 16);
		} else {
			ldc(&L.vbits)[(size_t)tile * 256 + 64 * wave + lane] = (uint8_t)v;
			const uint32_t n = wave_sum((uint32_t)__builtin_popcount(v));
			if (lane == 0)
				ldc(&L.tcount)[(size_t)tile * 4 + wave] = n;
			count += n;
		}
	}
	if (lane == 0)
		s_wcount[wave] = count;
	__syncthreads();
	const uint32_t n = s_wcount[0] + s_wcount[1] + s_wcount[2] + s_wcount[3];
	__syncthreads();                /* s_wcount is reused by the next call */
	return n;
}

template <int NS>
__global__ __launch_bounds__(256) void vg_count_kernel(const vg_prog P, const vg_launch L)
{
	extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
	const uint32_t g = blockIdx.x;
	const uint32_t t0 = g * L.tiles_per_wg;
	const uint32_t t1 = min(t0 + L.tiles_per_wg, L.num_tiles);
	if (g == 0 && threadIdx.x == 0)
		L.ctrl->overflow = 0;           /* the write pass that follows sets it */
	const uint32_t n = count_tiles<NS>(P, L, t0, t1, lds);
	if (threadIdx.x == 0)
		L.wgcount[g] = n;
}

/* ----------------------------------------------------------- write pass */
/* A workgroup of 4/WS waves works on one tile at a time; wave w takes
 * strips WS*w .. WS*w+WS-1 of it (lane l owns rows 256s + 4l .. +3 of
 * strip s).  Workgroup g owns tiles [g*wtiles_per_wg, (g+1)*wtiles_per_wg)
 * and starts its output after the rows selected before its first tile.
 * The per-strip counts of 64 tiles at a time are staged in LDS with their
 * running offsets; a wave visits only the tiles where it has selected
 * rows, loads what the result needs (only in lanes holding a selected row),
 * runs the micro-ops computing result registers, ranks its rows (strip by
 * strip, each in row order) and writes its contiguous slice of every
 * result column.  LDS: nslot * T words of registers, then 256*WS row
 * numbers per wave. */
#define VG_CHUNK 64
#ifndef VG_WS
#define VG_WS 1                               /* strips per write-pass wave */
#endif
#define VG_WRITE_THREADS (256 / VG_WS)

/* rows selected in this wave's strips of staged tile i */
__device__ __forceinline__ uint32_t wave_count(const uint32_t (*cnt)[4], uint32_t i, uint32_t wave)
{
	uint32_t n = 0;
#pragma unroll
	for (int s = 0; s < VG_WS; ++s)
		n += cnt[i][VG_WS * wave + s];
	return uni(n);
}

/* first tile index >= i of the chunk where wave `wave` has selected rows */
__device__ __forceinline__ uint32_t next_work(const uint32_t (*cnt)[4], uint32_t i, uint32_t nt,
	uint32_t wave)
{
	while (i < nt && wave_count(cnt, i, wave) == 0)
		++i;
	return i;
}

/* this lane's selected-row bits of a tile: strip WS*w+s in bits 4s..4s+3 */
__device__ __forceinline__ uint32_t load_bits(const vg_launch &L, uint32_t tile, uint32_t wave,
	uint32_t lane)
{
	const uint8_t *v = ldc(&L.vbits) + (size_t)tile * 256 + 64 * VG_WS * wave + lane;
	uint32_t b = 0;
#pragma unroll
	for (int s = 0; s < VG_WS; ++s)
		b |= (uint32_t)v[64 * s] << (4 * s);
	return b;
}

/* Point the cursor at `tile` and start loading what the write pass needs
 * for this lane's rows with selected bits; the caller waits (wait_vmem)
 * before the register file is read. */
__device__ __forceinline__ void issue_tile(const vg_prog &P, const vg_launch &L, cursor &C,
	uint32_t tile, uint32_t lane, const strips<VG_WS> &S, uint32_t bits, uint32_t *regs)
{
	cursor_seek(C, P, L, tile, lane);
	const uint32_t row0 = (tile - C.first) * VG_T;
	const uint32_t left = C.rows - row0;
	const uint32_t nrows = left < VG_T ? left : VG_T;
	if (bits)
		load_cols<VG_WS, VG_T>(P, C, row0, S, nrows, VG_LOAD_WRITE, regs, false);
}

/* Write n result rows of one wave: result row at+k comes from tile row
 * srcrow[k] (register slots, a column of the tile's tablet, or a
 * constant), lanes taking every 64th row. */
__device__ __forceinline__ void write_rows(const vg_prog &P, const vg_launch &L, const cursor &Cc,
	uint32_t row0, const uint32_t *regs, const uint16_t *srcrow, uint64_t at, uint32_t n, uint32_t lane)
{
	constexpr uint32_t T = VG_T;
	for (uint32_t j = 0; j < ldc(&P.nout); ++j) {
		const vg_out o = ldc_struct(&P.out[j]);
		char *dst = (char *)ldc(&L.out[j]);
		if (o.kind == VG_OUT_SLOT) {
			const uint32_t *lo = regs + o.slot * T, *hi = regs + o.slot_hi * T;
			if (o.width == 8) {
				for (uint32_t k = lane; k < n; k += 64) {
					const uint32_t s = srcrow[k];
					((uint64_t *)dst)[at + k] = ((uint64_t)hi[s] << 32) | lo[s];
				}
			} else if (o.width == 4) {
				for (uint32_t k = lane; k < n; k += 64)
					((uint32_t *)dst)[at + k] = lo[srcrow[k]];
			} else {
				for (uint32_t k = lane; k < n; k += 64)
					((uint8_t *)dst)[at + k] = (uint8_t)lo[srcrow[k]];
			}
		} else if (o.kind == VG_OUT_COLUMN) {
			const vg_col cc = cursor_col(Cc, o.col);
			const char *src = cursor_base(Cc, o.col) + (size_t)row0 * cc.width;
			if (o.width == 8) {
				for (uint32_t k = lane; k < n; k += 64)
					((uint64_t *)dst)[at + k] = gptr<uint64_t>(src)[srcrow[k]];
			} else if (o.width == 4) {
				for (uint32_t k = lane; k < n; k += 64)
					((uint32_t *)dst)[at + k] = gptr<uint32_t>(src)[srcrow[k]];
			} else {
				for (uint32_t k = lane; k < n; k += 64)
					((uint8_t *)dst)[at + k] = gptr<uint8_t>(src)[srcrow[k]];
			}
		} else {
			if (o.width == 8) {
				for (uint32_t k = lane; k < n; k += 64)
					((uint64_t *)dst)[at + k] = o.imm;
			} else if (o.width == 4) {
				for (uint32_t k = lane; k < n; k += 64)
					((uint32_t *)dst)[at + k] = (uint32_t)o.imm;
			} else {
				for (uint32_t k = lane; k < n; k += 64)
					((uint8_t *)dst)[at + k] = (uint8_t)o.imm;
			}
		}
	}
}

/* Write the selected rows of tiles [t0, t1) of a launch, the first at
 * output row `base`; returns the output row after the last. */
__device__ __forceinline__ uint64_t write_tiles(const vg_prog &P, const vg_launch &L, uint32_t t0,
	uint32_t t1, uint64_t base, uint32_t *lds)
{
	constexpr uint32_t T = VG_T;
	__shared__ uint32_t s_cnt[VG_CHUNK][4];
	__shared__ uint64_t s_off[VG_CHUNK];
	const uint32_t tid = threadIdx.x, lane = tid & 63, wave = uni(tid >> 6);
	uint32_t *regs = lds;
	uint16_t *srcrow = (uint16_t *)(lds + ldc(&P.nslot) * T) + wave * 256 * VG_WS;
	const strips<VG_WS> S{wave * 256u * VG_WS, 256u, lane};
	uint64_t chunk_base = base;
	cursor C;
	cursor_init(C, P, lane);

	for (uint32_t c0 = t0; c0 < t1; c0 += VG_CHUNK) {
		const uint32_t nt = min((uint32_t)VG_CHUNK, t1 - c0);
		if (wave == 0) {
			/* lane i: counts of tile c0+i and its offset inside the chunk */
			uint32_t c[4] = {0, 0, 0, 0};
			if (lane < nt) {
				const u32x4 v = *(const u32x4 *)(ldc(&L.tcount) + (size_t)(c0 + lane) * 4);
				c[0] = v.x; c[1] = v.y; c[2] = v.z; c[3] = v.w;
			}
			const uint32_t tot = c[0] + c[1] + c[2] + c[3];
			uint32_t incl = tot;
#pragma unroll
			for (int off = 1; off < 64; off <<= 1) {
				const uint32_t y = __shfl_up(incl, off, 64);
				if (lane >= (uint32_t)off)
					incl += y;
			}
			s_cnt[lane][0] = c[0]; s_cnt[lane][1] = c[1]; s_cnt[lane][2] = c[2]; s_cnt[lane][3] = c[3];
			s_off[lane] = chunk_base + (incl - tot);
		}
		__syncthreads();
		/* The tiles of the chunk where this wave has selected rows, in
		 * order.  The next one's validity bits are fetched while this one
		 * waits for its columns, and, when no output reads a loaded
		 * column's slot, the next one's columns load while this one's
		 * results are written. */
		uint32_t i = next_work(s_cnt, 0, nt, wave);
		uint32_t nbits = 0;
		if (i < nt) {
			nbits = load_bits(L, c0 + i, wave, lane);
			issue_tile(P, L, C, c0 + i, lane, S, nbits, regs);
		}
		while (i < nt) {
			uint32_t scnt[VG_WS], wave_off = 0;
#pragma unroll
			for (int s = 0; s < VG_WS; ++s)
				scnt[s] = uni(s_cnt[i][VG_WS * wave + s]);
			for (uint32_t q = 0; q < VG_WS * wave; ++q)
				wave_off += s_cnt[i][q];
			wave_off = uni(wave_off);
			const uint32_t wave_tot = wave_count(s_cnt, i, wave);
			const uint32_t tile = c0 + i;
			const uint32_t mybits = nbits;
			const cursor Cc = C;            /* this tile's tablet, for the gathers */
			const uint32_t inext = next_work(s_cnt, i + 1, nt, wave);
			if (inext < nt)
				nbits = load_bits(L, c0 + inext, wave, lane);
			wait_vmem();                    /* this tile's columns, the next bits */
			const uint32_t row0 = (tile - Cc.first) * T;
			rowstate RS;
			rows_init(RS, mybits);
			if (ldc(&P.wfull)) {
				/* rerun everything: the selected rows take the same
				 * branches as in the count pass, so masked writes see the
				 * same active rows */
				run_uops<VG_WS, T>(P, 0, ldc(&P.nuop), ~0ull, regs, S, RS);
			} else if (ldc(&P.wneed)) {
				run_uops<VG_WS, T>(P, 0, ldc(&P.nuop), ldc(&P.wneed), regs, S, RS);
			}
			/* the loaded columns' slots are free from here on */
			const bool early = ldc(&P.wsafe) && inext < nt;
			if (early)
				issue_tile(P, L, C, c0 + inext, lane, S, nbits, regs);
			/* ranks: strip by strip, each in row order */
			uint32_t sbase = 0;
#pragma unroll
			for (int s = 0; s < VG_WS; ++s) {
				uint32_t pos = sbase;
#pragma unroll
				for (int r = 0; r < 4; ++r)
					pos += rank_below(__ballot((mybits >> (4 * s + r)) & 1u));
#pragma unroll
				for (int r = 0; r < 4; ++r)
					if ((mybits >> (4 * s + r)) & 1u)
						srcrow[pos++] = (uint16_t)(S.row(s) + r);
				sbase += scnt[s];
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

			const uint64_t at = s_off[i] + wave_off;
			const uint64_t room = ldc(&L.capacity) > at ? ldc(&L.capacity) - at : 0;
			const uint32_t n = (uint64_t)wave_tot < room ? wave_tot : (uint32_t)room;
			if (n < wave_tot && lane == 0)
				atomicOr(&ldc(&L.ctrl)->overflow, 1u);
			write_rows(P, L, Cc, row0, regs, srcrow, at, n, lane);
			/* srcrow and the register rows are rewritten by the next tile */
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			if (!early && inext < nt)
				issue_tile(P, L, C, c0 + inext, lane, S, nbits, regs);
			i = inext;
		}
		{
			const uint32_t lastc = s_cnt[nt - 1][0] + s_cnt[nt - 1][1] + s_cnt[nt - 1][2] + s_cnt[nt - 1][3];
			chunk_base = s_off[nt - 1] + lastc;
		}
		__syncthreads();
	}
	return chunk_base;
}

/* A run launched with an outcome pointer (a query run alone while the
 * host polls): one thread after the write pass stores the row count and
 * overflow bits into pinned host memory, `done` last.  The launch boundary
 * orders it after every result store of the write pass. */
__global__ __launch_bounds__(64) void vg_publish_kernel(const vg_ctrl *c, vg_bresult *h)
{
	if (threadIdx.x != 0)
		return;
	vg_bresult r;
	r.total = c->total;
	r.flags = c->overflow;
	r.t_done = wall_clock64();
	r.t_start = 0;
	r.done = 0;
	*h = r;
	__threadfence_system();
	__hip_atomic_store(&h->done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ __launch_bounds__(VG_WRITE_THREADS) void vg_write_kernel(const vg_prog P, const vg_launch L)
{
	extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
	__shared__ uint64_t s_base;
	const uint32_t tid = threadIdx.x, lane = tid & 63, wave = uni(tid >> 6);
	const uint32_t g = blockIdx.x;
	const uint32_t t0 = g * L.wtiles_per_wg;
	const uint32_t t1 = min(t0 + L.wtiles_per_wg, L.num_tiles);

	{
		/* offset of this workgroup: the rows selected before tile t0 --
		 * whole count-pass workgroups, then the tiles of the one t0 is
		 * in; all threads add, eight independent loads in flight each */
		__shared__ uint64_t s_part[VG_WRITE_THREADS / 64];
		const uint32_t cw = t0 / L.tiles_per_wg;
		uint64_t base = 0;
		for (uint32_t k0 = 0; k0 < cw; k0 += 8 * VG_WRITE_THREADS) {
			uint32_t v[8];
#pragma unroll
			for (int j = 0; j < 8; ++j) {
				const uint32_t k = k0 + j * VG_WRITE_THREADS + tid;
				v[j] = k < cw ? L.wgcount[k] : 0u;
			}
#pragma unroll
			for (int j = 0; j < 8; ++j)
				base += v[j];
		}
		for (uint32_t t = cw * L.tiles_per_wg + tid; t < t0; t += VG_WRITE_THREADS) {
			const u32x4 v = *(const u32x4 *)(L.tcount + (size_t)t * 4);
			base += (uint64_t)v.x + v.y + v.z + v.w;
		}
#pragma unroll
		for (int off = 32; off > 0; off >>= 1)
			base += __shfl_xor(base, off, 64);
		if (lane == 0)
			s_part[wave] = base;
		__syncthreads();
		if (tid == 0) {
			uint64_t sum = 0;
#pragma unroll
			for (uint32_t w = 0; w < VG_WRITE_THREADS / 64; ++w)
				sum += s_part[w];
			s_base = sum;
		}
	}
	__syncthreads();
	const uint64_t end = write_tiles(P, L, t0, t1, s_base, lds);
	if (g == gridDim.x - 1 && tid == 0)
		L.ctrl->total = end;
}

/* ------------------------------------------------------ batched queries */
/* Persistent launches run the items of many queries (each with its own
 * program, tablets and outputs): a count launch (4 waves share a tile,
 * as the count pass), a scan launch turning each query's item counts into
 * output offsets, a write launch (as the write pass) and a finish launch
 * gathering each query's overflow bits and finish time.  Within a launch
 * no workgroup waits for another, so items are dealt out statically:
 * workgroup g takes items g, g + grid, ...  (A single atomic work counter
 * serialises device-wide, about 75 ns a grab, which cost more than the
 * balance it bought.)  The single-pass launch further down does wait on
 * other workgroups and takes its items from a counter in list order. */

/* clear the per-query state and the item counters; stamp the start */
__global__ __launch_bounds__(256) void vg_batch_reset(const vg_batch_args A)
{
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i < A.nquery) {
		vg_bquery z;
		memset(&z, 0, sizeof z);
		A.q[i] = z;
		vg_ctrl *c = A.launch[i].ctrl;
		c->total = 0;
		c->overflow = 0;
		vg_bresult r;
		memset(&r, 0, sizeof r);
		A.hres[i] = r;
	}
	if (i < A.nhead)
		A.head[i].t_start = wall_clock64();
	for (uint64_t j = i; j < A.nstatus; j += (uint64_t)gridDim.x * 256)
		A.status[j] = 0;
}

/* Run by the workgroup finishing a query's last item: the query's row
 * count and overflow bits into its batch state (one copy brings every
 * query's back), and the time. */
__device__ __forceinline__ void finish_query(const vg_batch_args &A, uint32_t q)
{
	vg_ctrl *c = ldc(&A.launch[q].ctrl);
	vg_bresult r;
	r.total = __hip_atomic_load(&c->total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	r.flags = __hip_atomic_load(&c->overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	r.t_done = wall_clock64();
	r.t_start = ldc(&A.head)->t_start;
	r.done = 0;
	A.q[q].total = r.total;
	A.q[q].flags = r.flags;
	A.q[q].t_done = r.t_done;
	A.hres[q] = r;
	__threadfence_system();
	__hip_atomic_store(&A.hres[q].done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

/* The single pass's next item: each query's tiles are taken in order
 * through the query's own counter, so workgroups on different queries do
 * not contend on one device-wide counter.  A workgroup stays with a query
 * (`cur`, an index into the group's queries; thread 0's copy is the one
 * used) until its tiles are all taken, then moves on, and gets query ~0u
 * when a whole round over the group's queries found none.  Within a query
 * tiles are still taken in order by running workgroups, which the
 * look-back needs. */
__device__ __forceinline__ void next_query_item(const vg_batch_args &A, uint32_t &cur, vg_item *s_item,
	uint32_t &q, uint32_t &t0, uint32_t &t1, uint32_t &k)
{
	if (threadIdx.x == 0) {
		uint32_t iq = ~0u, i0 = 0, i1 = 0, ik = 0;
		const uint32_t ngq = A.ngq;
		for (uint32_t tries = 0; tries < ngq; ++tries) {
			const uint32_t qq = ldc(&A.gq[cur]);
			const uint32_t i = atomicAdd(&A.q[qq].next, 1u);
			if (i < ldc(&A.info[qq].nitem)) {
				const u32x4 v = *(const u32x4 *)(A.witems + ldc(&A.info[qq].wstart) + i);
				iq = v.x; i0 = v.y; i1 = v.z; ik = v.w;
				break;
			}
			cur = cur + 1 == ngq ? 0 : cur + 1;
		}
		s_item->query = iq;
		s_item->t0 = i0;
		s_item->t1 = i1;
		s_item->k = ik;
	}
	__syncthreads();
	q = uni(s_item->query);
	t0 = uni(s_item->t0);
	t1 = uni(s_item->t1);
	k = uni(s_item->k);
	__syncthreads();                /* s_item is rewritten by the next fetch */
}

/* Count items, statically: workgroup g takes items g, g + grid, ...  Each
 * item's count goes to icount; nothing waits on another workgroup. */
__global__ __launch_bounds__(256) void vg_batch_count_kernel(const vg_batch_args A)
{
	extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
	for (uint32_t i = blockIdx.x; i < A.nitems; i += gridDim.x) {
		const u32x4 it = ldc((const u32x4 *)(A.items + i));
		const uint32_t q = it.x, t0 = it.y, t1 = it.z, k = it.w;
		const uint32_t n = count_tiles<1>(A.progs[q], A.launch[q], t0, t1, lds);
		if (threadIdx.x == 0)
			A.icount[ldc(&A.info[q].ioff) + k] = n;
	}
}

/* Per query of the launch's group (one wave each): exclusive scan of its
 * items' counts into output offsets, and its row count. */
__global__ __launch_bounds__(64) void vg_batch_scan_kernel(const vg_batch_args A)
{
	const uint32_t lane = threadIdx.x;
	if (blockIdx.x >= A.ngq)
		return;
	const uint32_t q = ldc(&A.gq[blockIdx.x]);
	const uint32_t nitem = ldc(&A.info[q].nitem), ioff = ldc(&A.info[q].ioff);
	uint64_t carry = 0;
	for (uint32_t c = 0; c < nitem; c += 64) {
		const uint32_t v = c + lane < nitem ? A.icount[ioff + c + lane] : 0u;
		uint64_t incl = v;
#pragma unroll
		for (int off = 1; off < 64; off <<= 1) {
			const uint64_t y = __shfl_up(incl, off, 64);
			if (lane >= (uint32_t)off)
				incl += y;
		}
		if (c + lane < nitem)
			A.ibase[ioff + c + lane] = carry + incl - v;
		carry += __shfl(incl, 63, 64);
	}
	if (lane == 0) {
		A.q[q].total = carry;
		ldc(&A.launch[q].ctrl)->total = carry;
	}
}

/* Write items, statically like the count items; each stamps its finish
 * time (device clock) for the per-query latency. */
__global__ __launch_bounds__(VG_WRITE_THREADS) void vg_batch_write_kernel(const vg_batch_args A)
{
	extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
	for (uint32_t i = blockIdx.x; i < A.nitems; i += gridDim.x) {
		const u32x4 it = ldc((const u32x4 *)(A.witems + i));
		const uint32_t q = it.x, t0 = it.y, t1 = it.z, k = it.w;
		write_tiles(A.progs[q], A.launch[q], t0, t1, A.ibase[ldc(&A.info[q].ioff) + k], lds);
		if (threadIdx.x == 0)
			A.stamp[i] = wall_clock64();
		/* Keep this barrier: without one between an item's thread-0
		 * block and the next item, the compiler may merge the two across
		 * the loop's back edge. */
		__syncthreads();
	}
}

/* Per query of the group (one wave each): overflow bits from its control
 * block and its finish time, the latest stamp of its write items (which
 * are consecutive in the write list from info.wstart). */
__global__ __launch_bounds__(64) void vg_batch_finish_kernel(const vg_batch_args A)
{
	const uint32_t lane = threadIdx.x;
	if (blockIdx.x >= A.ngq)
		return;
	const uint32_t q = ldc(&A.gq[blockIdx.x]);
	const uint32_t nitem = ldc(&A.info[q].nitem), w0 = ldc(&A.info[q].wstart);
	uint64_t t = 0;
	for (uint32_t c = lane; c < nitem; c += 64)
		t = max(t, A.stamp[w0 + c]);
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		const uint64_t y = __shfl_xor(t, off, 64);
		t = t > y ? t : y;
	}
	if (lane == 0) {
		vg_bresult r;
		r.total = A.q[q].total;
		r.flags = ldc(&A.launch[q].ctrl)->overflow;
		r.t_done = t;
		r.t_start = ldc(&A.head)->t_start;
		r.done = 0;
		A.q[q].flags = r.flags;
		A.q[q].t_done = t;
		A.hres[q] = r;
		__threadfence_system();
		__hip_atomic_store(&A.hres[q].done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
	}
}

/* 64-bit sum over the wave */
__device__ __forceinline__ uint64_t wave_sum64(uint64_t x)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		x += __shfl_xor(x, off, 64);
	return x;
}

/* Decoupled look-back (run by one whole wave): publish tile `tile`'s count
 * n, add up the counts of the tiles before it -- the nearest published
 * running total plus the single counts after it -- and publish the running
 * total through this tile.  Returns the rows selected before the tile, or
 * ~0 when a predecessor stayed unpublished for spin_limit polls. */
__device__ __forceinline__ uint64_t lookback(uint64_t *st, uint32_t tile, uint32_t n, uint32_t lane,
	uint32_t spin_limit)
{
	if (tile == 0) {
		if (lane == 0)
			__hip_atomic_store(&st[0], VG_ST_PRE | n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		return 0;
	}
	if (lane == 0)
		__hip_atomic_store(&st[tile], VG_ST_AGG | n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	uint64_t sum = 0;
	int64_t j0 = (int64_t)tile - 1;         /* lane l looks at tile j0 - l */
	uint32_t polls = 0;
	for (;;) {
		const int64_t j = j0 - (int64_t)lane;
		const uint64_t v = j >= 0 ? __hip_atomic_load(&st[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
		                          : VG_ST_PRE;          /* before tile 0: total 0 */
		const uint32_t f = (uint32_t)(v >> 62);
		const uint64_t pm = __ballot(f == 2u), nm = __ballot(f == 0u);
		const uint64_t first = pm & (~pm + 1);          /* nearest running total */
		const uint64_t upto = pm ? (first | (first - 1)) : ~0ull;
		if (nm & upto) {
			/* a nearer tile has not published yet */
			if (++polls >= spin_limit)
				return ~0ull;
			__builtin_amdgcn_s_sleep(1);
			continue;
		}
		sum += wave_sum64(((upto >> lane) & 1ull) ? (v & VG_ST_VAL) : 0ull);
		if (pm)
			break;
		j0 -= 64;
	}
	if (lane == 0)
		__hip_atomic_store(&st[tile], VG_ST_PRE | (sum + n), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	return sum;
}

/* Single pass over chunk items (each query's chunks of tiles listed in
 * order, taken in order through the query's own counter): the workgroup
 * counts the chunk as the count pass does (validity bytes and per-strip
 * counts into the query's two-pass arrays), publishes the chunk's count at
 * once, takes the chunk's output offset by look-back over the query's
 * earlier chunks (one status word per chunk), and writes the chunk's
 * selected rows there as the write pass does, reading the filter columns
 * again from L2.  The output capacity is the caller's estimate; past it
 * rows are dropped and VG_OVF_CAPACITY set (the exact count is still
 * produced).  The workgroup finishing a query's last chunk stores the
 * query's outcome at once. */
__global__ __launch_bounds__(256) void vg_batch_single_kernel(const vg_batch_args A)
{
	extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
	__shared__ vg_item s_item;
	__shared__ uint64_t s_pre;
	const uint32_t tid = threadIdx.x, lane = tid & 63, wave = uni(tid >> 6);
	uint32_t cur = A.ngq ? blockIdx.x % A.ngq : 0;
	for (;;) {
		uint32_t q, t0, t1, k;
		next_query_item(A, cur, &s_item, q, t0, t1, k);
		if (q == ~0u)
			break;
		const vg_prog &P = A.progs[q];
		const vg_launch &L = A.launch[q];
		uint64_t *st = A.status + ldc(&A.info[q].soff);
		const uint32_t nitem = ldc(&A.info[q].nitem);
		const uint32_t n = count_tiles<1>(P, L, t0, t1, lds);
		if (wave == 0) {
			const uint64_t pre = lookback(st, k, n, lane, A.spin_limit);
			if (lane == 0)
				s_pre = pre;
		}
		__syncthreads();
		const uint64_t pre = s_pre;
		if (pre == ~0ull) {
			if (tid == 0)
				atomicOr(&ldc(&L.ctrl)->overflow, VG_OVF_WAIT);
		} else if (n != 0) {
			write_tiles(P, L, t0, t1, pre, lds);
		}
		if (pre != ~0ull && k + 1 == nitem && tid == 0)
			ldc(&L.ctrl)->total = pre + n;
		__syncthreads();            /* the write phase's stores before the count below */
		if (tid == 0 && __hip_atomic_fetch_add(&A.q[q].wdone, 1u, __ATOMIC_ACQ_REL,
				__HIP_MEMORY_SCOPE_AGENT) + 1 == nitem)
			finish_query(A, q);
		__syncthreads();            /* see the write kernel: keeps the fetch in the loop */
	}
}

/* --------------------------------------------------------------- launch */
/* the count pass takes a whole tile per wave when the filter's register
 * file is this small */
#ifndef VG_COUNT_WIDE_SLOTS
#define VG_COUNT_WIDE_SLOTS 2u
#endif

static bool count_wide(const vg_prog *P)
{
	return P->nslot_count <= VG_COUNT_WIDE_SLOTS;
}

static size_t count_lds_bytes(const vg_prog *P)
{
	return (size_t)(count_wide(P) ? 4 : 1) * P->nslot_count * VG_T * 4;
}

static size_t write_lds_bytes(const vg_prog *P)
{
	return (size_t)P->nslot * VG_T * 4 + 4 * 256 * 2;
}

/* the gfx950 LDS is 160 KB per CU; keep room for the static arrays */
#define VG_LDS_LIMIT (150u * 1024u)

extern "C" size_t vg_kernel_lds_bytes(const vg_prog *P)
{
	const size_t a = count_lds_bytes(P), b = write_lds_bytes(P);
	return a > b ? a : b;
}

static int num_cus()
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
	return cus;
}

static void allow_lds()
{
	static bool done = false;
	if (!done) {
		(void)hipFuncSetAttribute((const void *)vg_count_kernel<4>, hipFuncAttributeMaxDynamicSharedMemorySize, VG_LDS_LIMIT);
		(void)hipFuncSetAttribute((const void *)vg_count_kernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize, VG_LDS_LIMIT);
		(void)hipFuncSetAttribute((const void *)vg_write_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, VG_LDS_LIMIT);
		done = true;
	}
}

/* Workgroups per pass: every CU filled at the pass's occupancy, never
 * more than there are tiles. */
/* workgroups of a pass kernel (0 count wide, 1 count, 2 write) that fit
 * on the device with `lds` bytes each, asked once per (kernel, lds) */
static uint32_t pass_fit(int which, size_t lds)
{
	static std::unordered_map<uint64_t, uint32_t> memo;
	const uint64_t key = ((uint64_t)which << 32) | (uint64_t)lds;
	auto it = memo.find(key);
	if (it != memo.end())
		return it->second;
	allow_lds();
	int per = 0;
	const hipError_t e = which == 0
		? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, vg_count_kernel<4>, 256, lds)
		: which == 1
		? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, vg_count_kernel<1>, 256, lds)
		: hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, vg_write_kernel, VG_WRITE_THREADS, lds);
	if (e != hipSuccess || per < 1)
		per = 1;
	const uint32_t g = (uint32_t)num_cus() * (uint32_t)per;
	memo[key] = g;
	return g;
}

static uint32_t count_grid(const vg_prog *P, uint32_t num_tiles)
{
	const uint32_t g = pass_fit(count_wide(P) ? 0 : 1, count_lds_bytes(P));
	return g < num_tiles ? g : num_tiles;
}

static uint32_t write_grid(const vg_prog *P, uint32_t num_tiles)
{
	const uint32_t g = pass_fit(2, write_lds_bytes(P));
	return g < num_tiles ? g : num_tiles;
}

/* split num_tiles into runs of *per tiles for at most `want` workgroups;
 * returns the number of workgroups */
static uint32_t split(uint32_t num_tiles, uint32_t want, uint32_t *per)
{
	*per = (num_tiles + want - 1) / want;
	return (num_tiles + *per - 1) / *per;
}

/* Number of count-pass workgroups of a launch (the size of the
 * per-workgroup count array). */
extern "C" uint32_t vg_kernel_grid(const vg_prog *P, uint32_t num_tiles)
{
	if (num_tiles == 0)
		return 0;
	uint32_t per;
	return split(num_tiles, count_grid(P, num_tiles), &per);
}

/* 0 if the program's register file fits the LDS of both passes */
extern "C" int vg_kernel_check(const vg_prog *P)
{
	return vg_kernel_lds_bytes(P) > VG_LDS_LIMIT ? -3 : 0;
}

/* Count pass then write pass on `stream`; `mid`, if given, is recorded
 * between them (per-pass timing). */
extern "C" int vg_kernel_launch(const vg_prog *P, vg_launch *L, hipStream_t stream, hipEvent_t mid)
{
	if (L->num_tiles == 0)
		return 0;
	if (vg_kernel_check(P) != 0)
		return -3;
	const uint32_t grid = split(L->num_tiles, count_grid(P, L->num_tiles), &L->tiles_per_wg);
	const uint32_t wgrid = split(L->num_tiles, write_grid(P, L->num_tiles), &L->wtiles_per_wg);
	if (count_wide(P))
		hipLaunchKernelGGL(vg_count_kernel<4>, dim3(grid), dim3(256), count_lds_bytes(P), stream, *P, *L);
	else
		hipLaunchKernelGGL(vg_count_kernel<1>, dim3(grid), dim3(256), count_lds_bytes(P), stream, *P, *L);
	if (hipGetLastError() != hipSuccess)
		return -2;
	if (mid && hipEventRecord(mid, stream) != hipSuccess)
		return -2;
	hipLaunchKernelGGL(vg_write_kernel, dim3(wgrid), dim3(VG_WRITE_THREADS), write_lds_bytes(P), stream, *P, *L);
	if (L->hres)
		hipLaunchKernelGGL(vg_publish_kernel, dim3(1), dim3(64), 0, stream, (const vg_ctrl *)L->ctrl, L->hres);
	return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" uint32_t vg_kernel_tile_rows(void)
{
	return VG_T;
}

/* dynamic LDS of a program's batch count items (pass 0: the tile's
 * register file shared by the 4 waves), write items (pass 1) or
 * single-pass items (pass 2) */
extern "C" size_t vg_kernel_batch_lds(int pass, const vg_prog *P)
{
	return pass == 0 ? (size_t)P->nslot_count * VG_T * 4 : write_lds_bytes(P);
}

static void batch_attrs()
{
	static bool done = false;
	if (!done) {
		(void)hipFuncSetAttribute((const void *)vg_batch_count_kernel,
			hipFuncAttributeMaxDynamicSharedMemorySize, VG_LDS_LIMIT);
		(void)hipFuncSetAttribute((const void *)vg_batch_write_kernel,
			hipFuncAttributeMaxDynamicSharedMemorySize, VG_LDS_LIMIT);
		(void)hipFuncSetAttribute((const void *)vg_batch_single_kernel,
			hipFuncAttributeMaxDynamicSharedMemorySize, VG_LDS_LIMIT);
		done = true;
	}
}

/* Workgroups of the batch count (pass 0) or write (pass 1) kernel that
 * fit on the device at once with `lds` bytes of dynamic LDS each. */
extern "C" uint32_t vg_kernel_batch_grid(int pass, size_t lds)
{
	/* the occupancy query is slow next to a batch build: remember answers */
	static std::unordered_map<uint64_t, uint32_t> memo;
	const uint64_t key = ((uint64_t)pass << 32) | (uint64_t)lds;
	auto it = memo.find(key);
	if (it != memo.end())
		return it->second;
	batch_attrs();
	int per = 0;
	const hipError_t e = pass == 0
		? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, vg_batch_count_kernel, 256, lds)
		: pass == 1
		? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, vg_batch_write_kernel, VG_WRITE_THREADS, lds)
		: hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, vg_batch_single_kernel, 256, lds);
	if (e != hipSuccess || per < 1)
		per = 1;
	const uint32_t g = (uint32_t)num_cus() * (uint32_t)per;
	memo[key] = g;
	return g;
}

/* Clear the state of every query and of A->nhead item-list heads. */
extern "C" int vg_kernel_batch_reset(const vg_batch_args *A, hipStream_t stream)
{
	uint64_t m = A->nquery > A->nhead ? A->nquery : A->nhead;
	m = m > A->nstatus ? m : (A->nstatus < 256u * 1024u ? A->nstatus : 256u * 1024u);
	const uint32_t rgrid = (uint32_t)((m + 255) / 256);
	hipLaunchKernelGGL(vg_batch_reset, dim3(rgrid ? rgrid : 1), dim3(256), 0, stream, *A);
	return hipGetLastError() == hipSuccess ? 0 : -2;
}

/* The count launch, then the write launch, over A's item lists. */
extern "C" int vg_kernel_batch_passes(const vg_batch_args *A, uint32_t cgrid, size_t clds, uint32_t wgrid,
	size_t wlds, hipStream_t stream)
{
	if (clds > VG_LDS_LIMIT || wlds > VG_LDS_LIMIT)
		return -3;
	batch_attrs();
	hipLaunchKernelGGL(vg_batch_count_kernel, dim3(cgrid), dim3(256), clds, stream, *A);
	if (hipGetLastError() != hipSuccess)
		return -2;
	hipLaunchKernelGGL(vg_batch_scan_kernel, dim3(A->ngq ? A->ngq : 1), dim3(64), 0, stream, *A);
	if (hipGetLastError() != hipSuccess)
		return -2;
	hipLaunchKernelGGL(vg_batch_write_kernel, dim3(wgrid), dim3(VG_WRITE_THREADS), wlds, stream, *A);
	if (hipGetLastError() != hipSuccess)
		return -2;
	hipLaunchKernelGGL(vg_batch_finish_kernel, dim3(A->ngq ? A->ngq : 1), dim3(64), 0, stream, *A);
	return hipGetLastError() == hipSuccess ? 0 : -2;
}

/* The single-pass launch over A's (write) item list. */
extern "C" int vg_kernel_batch_single(const vg_batch_args *A, uint32_t grid, size_t lds, hipStream_t stream)
{
	if (lds > VG_LDS_LIMIT)
		return -3;
	batch_attrs();
	hipLaunchKernelGGL(vg_batch_single_kernel, dim3(grid), dim3(256), lds, stream, *A);
	return hipGetLastError() == hipSuccess ? 0 : -2;
}
