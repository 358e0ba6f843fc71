/*
 * Host stand-ins for the device built-ins the specialised kernels use, so
 * that a program's generated kernel source compiles for the CPU and runs
 * one workgroup at a time (tests/kernel_cpu.py).  A lane is a coroutine, a
 * wave 64 of them, a workgroup 256; a cross-lane operation writes each
 * lane's value to the wave's exchange slots between two wave barriers.
 * Buffer loads check every dword against the resource's size as the
 * hardware does, so the kernels' masked and out-of-range loads read 0.
 * A barrier that some lanes can no longer reach (a cross-lane operation in
 * divergent code, a lost wait) aborts the run.
 *
 * Test infrastructure only: nothing of the product includes it.
 */
#ifndef KERNEL_CPU_SHIM_H
#define KERNEL_CPU_SHIM_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ucontext.h>

#define __HIPCC_RTC__ 1
#define __device__
#define __forceinline__ inline
#define __global__
#define __launch_bounds__(...)
#define __shared__ static

struct kc_dim {
	unsigned x, y, z;
};
extern kc_dim threadIdx;
extern kc_dim gridDim;

/* The workgroup's 256 lanes are coroutines of one host thread: a lane runs
 * until it waits at a barrier; the scheduler then runs the next lane that
 * can run, and releases a barrier's lanes once all of them have arrived. */
struct kc_barrier {
	unsigned n = 0, count = 0;
	unsigned long gen = 0;
	void wait();
};

struct kc_block {
	kc_barrier all;
	kc_barrier wave[4];
	uint64_t slot[4][64];
};
extern kc_block *kc_blk;

static inline unsigned kc_lane() { return threadIdx.x & 63u; }
static inline unsigned kc_wavei() { return threadIdx.x >> 6; }

/* every lane's value of the wave, then f over them */
template <typename F> static inline uint64_t kc_cross(uint64_t v, F f)
{
	kc_block *b = kc_blk;
	const unsigned w = kc_wavei();
	b->slot[w][kc_lane()] = v;
	b->wave[w].wait();
	const uint64_t r = f(b->slot[w]);
	b->wave[w].wait();
	return r;
}

static inline unsigned __builtin_amdgcn_readfirstlane_kc(unsigned x)
{
	return (unsigned)kc_cross(x, [](const uint64_t *s) { return s[0]; });
}
#define __builtin_amdgcn_readfirstlane(x) __builtin_amdgcn_readfirstlane_kc((unsigned)(x))

template <typename T> static inline T __shfl_xor(T x, int o, int)
{
	uint64_t bits = 0;
	memcpy(&bits, &x, sizeof x);
	const unsigned lane = kc_lane();
	const uint64_t r = kc_cross(bits, [&](const uint64_t *s) { return s[lane ^ (unsigned)o]; });
	T out;
	memcpy(&out, &r, sizeof out);
	return out;
}

static inline unsigned long long __ballot(int pred)
{
	return kc_cross(pred ? 1u : 0u, [](const uint64_t *s) {
		uint64_t m = 0;
		for (unsigned i = 0; i < 64; i++)
			m |= s[i] << i;
		return m;
	});
}

static inline unsigned kc_mbcnt_lo(unsigned m, unsigned acc)
{
	const unsigned lane = kc_lane();
	const unsigned mask = lane >= 32 ? 0xFFFFFFFFu : (1u << lane) - 1u;
	return acc + (unsigned)__builtin_popcount(m & mask);
}
static inline unsigned kc_mbcnt_hi(unsigned m, unsigned acc)
{
	const unsigned lane = kc_lane();
	const unsigned mask = lane < 32 ? 0u : (lane == 32 ? 0u : (1u << (lane - 32)) - 1u);
	return acc + (unsigned)__builtin_popcount(m & mask);
}
#define __builtin_amdgcn_mbcnt_lo(m, a) kc_mbcnt_lo((unsigned)(m), (unsigned)(a))
#define __builtin_amdgcn_mbcnt_hi(m, a) kc_mbcnt_hi((unsigned)(m), (unsigned)(a))

#define __builtin_amdgcn_s_waitcnt(x) ((void)0)
#define __builtin_amdgcn_s_sleep(x) ((void)0)
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(__ATOMIC_SEQ_CST)
#define __builtin_amdgcn_wave_barrier() kc_blk->wave[kc_wavei()].wait()

static inline void __syncthreads()
{
	__atomic_thread_fence(__ATOMIC_SEQ_CST);
	kc_blk->all.wait();
	__atomic_thread_fence(__ATOMIC_SEQ_CST);
}

#define __HIP_MEMORY_SCOPE_WAVEFRONT 0
#define __HIP_MEMORY_SCOPE_WORKGROUP 0
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __HIP_MEMORY_SCOPE_SYSTEM 0
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_add(p, v, order, scope) __atomic_fetch_add((p), (v), __ATOMIC_SEQ_CST)
template <typename T, typename U> static inline T atomicAdd(T *p, U v) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
template <typename T, typename U> static inline T atomicOr(T *p, U v) { return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST); }
static inline unsigned long long wall_clock64() { return 0; }

static inline unsigned __float_as_uint(float f) { unsigned u; memcpy(&u, &f, 4); return u; }
static inline float __uint_as_float(unsigned u) { float f; memcpy(&f, &u, 4); return f; }
static inline double __longlong_as_double(long long v) { double d; memcpy(&d, &v, 8); return d; }
static inline long long __double_as_longlong(double d) { long long v; memcpy(&v, &d, 8); return v; }
template <typename T> static inline T min(T a, T b) { return b < a ? b : a; }

/* buffer resources: base and size in bytes; a dword (a byte for b8) past
 * the size reads 0 */
struct kc_rsrc {
	const char *base;
	uint32_t num;
};
#define __amdgpu_buffer_rsrc_t kc_rsrc
static inline kc_rsrc kc_make_rsrc(const void *p, int, int num, int) { return kc_rsrc{(const char *)p, (uint32_t)num}; }
#define __builtin_amdgcn_make_buffer_rsrc(p, stride, num, flags) kc_make_rsrc((p), (stride), (num), (flags))
typedef unsigned kc_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned kc_u32x2 __attribute__((ext_vector_type(2)));
static inline unsigned kc_dword(kc_rsrc r, uint64_t off)
{
	unsigned v = 0;
	if (off + 4 <= r.num)
		memcpy(&v, r.base + off, 4);
	return v;
}
static inline kc_u32x4 kc_load_b128(kc_rsrc r, unsigned off, int soff, int)
{
	const uint64_t o = (uint64_t)off + (unsigned)soff;
	kc_u32x4 v = {kc_dword(r, o), kc_dword(r, o + 4), kc_dword(r, o + 8), kc_dword(r, o + 12)};
	return v;
}
static inline kc_u32x2 kc_load_b64(kc_rsrc r, unsigned off, int soff, int)
{
	const uint64_t o = (uint64_t)off + (unsigned)soff;
	kc_u32x2 v = {kc_dword(r, o), kc_dword(r, o + 4)};
	return v;
}
static inline unsigned kc_load_b32(kc_rsrc r, unsigned off, int soff, int)
{
	return kc_dword(r, (uint64_t)off + (unsigned)soff);
}
static inline unsigned char kc_load_b8(kc_rsrc r, unsigned off, int soff, int)
{
	const uint64_t o = (uint64_t)off + (unsigned)soff;
	return o < r.num ? (unsigned char)r.base[o] : 0;
}
#define __builtin_amdgcn_raw_buffer_load_b128 kc_load_b128
#define __builtin_amdgcn_raw_buffer_load_b64 kc_load_b64
#define __builtin_amdgcn_raw_buffer_load_b32 kc_load_b32
#define __builtin_amdgcn_raw_buffer_load_b8 kc_load_b8

/* a lane's `size` bytes from src land at dst + lane * size */
static inline void kc_load_lds(const void *src, void *dst, int size, int, int)
{
	memcpy((char *)dst + kc_lane() * (unsigned)size, src, (size_t)size);
}
#define __builtin_amdgcn_global_load_lds kc_load_lds

#endif
