/*
 * The synthetic tables' values, one row at a time.  Compiled into the host
 * library (gen.c) and into the device table builder (gpu/gen.hip), so both
 * produce the same bits; both are compiled with -ffp-contract=off, and the
 * arithmetic is IEEE add, multiply, divide, sqrt, floor and conversions,
 * which round the same way on either side.
 *
 * Every value is a pure function of (seed, column, row): Philox4x32-10
 * (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
 * 1, 2, 3", SC'11) keyed by (seed, column) with counter (row, attempt),
 * and normals from the Marsaglia polar method with a logarithm written
 * here from IEEE add/multiply/divide only, so any machine (and numpy)
 * reproduces the tables bit for bit.
 */
#ifndef VG_GEN_MATH_H
#define VG_GEN_MATH_H

#include <stdint.h>

#ifdef __HIPCC__
#define VG_GEN_FN static __host__ __device__ inline
#else
#include <math.h>
#define VG_GEN_FN static inline
#endif

enum { VG_GC_UNIFORMI = 1, VG_GC_NORMALI5, VG_GC_NORMALI20, VG_GC_UNIFORMF, VG_GC_NORMALF5, VG_GC_NORMALF20 };

VG_GEN_FN void vg_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
	uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
	uint32_t k0 = key[0], k1 = key[1];
	for (int i = 0; i < 10; i++) {
		uint64_t p0 = (uint64_t)0xD2511F53u * c0;
		uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
		uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
		uint32_t n1 = (uint32_t)p1;
		uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
		uint32_t n3 = (uint32_t)p0;
		c0 = n0; c1 = n1; c2 = n2; c3 = n3;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* a double in [0,1) from 53 random bits */
VG_GEN_FN double vg_u53(uint32_t a, uint32_t b)
{
	return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* natural logarithm of x in (0, 1]: x = m * 2^e with m in [sqrt(1/2),
 * sqrt(2)), log m = 2 atanh(f), f = (m-1)/(m+1), |f| < 0.1716, by its odd
 * series to f^19 (error below 1e-17 relative). */
VG_GEN_FN double vg_log(double x)
{
	uint64_t bits;
	__builtin_memcpy(&bits, &x, 8);
	int e = (int)((bits >> 52) & 0x7FF) - 1023;
	bits = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
	double m;
	__builtin_memcpy(&m, &bits, 8);
	if (m > 1.4142135623730951) {
		m = m * 0.5;
		e = e + 1;
	}
	double f = (m - 1.0) / (m + 1.0);
	double f2 = f * f;
	double p = 1.0 / 19.0;
	p = p * f2 + 1.0 / 17.0;
	p = p * f2 + 1.0 / 15.0;
	p = p * f2 + 1.0 / 13.0;
	p = p * f2 + 1.0 / 11.0;
	p = p * f2 + 1.0 / 9.0;
	p = p * f2 + 1.0 / 7.0;
	p = p * f2 + 1.0 / 5.0;
	p = p * f2 + 1.0 / 3.0;
	p = p * f2 + 1.0;
	return (double)e * 0.6931471805599453 + 2.0 * f * p;
}

/* standard normal for (seed, column, row) */
VG_GEN_FN double vg_gen_normal(uint32_t seed, uint32_t column, uint64_t row)
{
	uint32_t key[2] = { seed, column };
	for (uint32_t attempt = 0;; attempt++) {
		uint32_t ctr[4] = { (uint32_t)row, (uint32_t)(row >> 32), attempt, 0x4E4F524Du };
		uint32_t w[4];
		vg_philox(ctr, key, w);
		double a = 2.0 * vg_u53(w[0], w[1]) - 1.0;
		double b = 2.0 * vg_u53(w[2], w[3]) - 1.0;
		double s = a * a + b * b;
		if (s >= 1.0 || s == 0.0)
			continue;
		return a * sqrt(-2.0 * vg_log(s) / s);
	}
}

/* uniform double in [0,1) for (seed, column, row) */
VG_GEN_FN double vg_gen_uniform(uint32_t seed, uint32_t column, uint64_t row)
{
	uint32_t key[2] = { seed, column };
	uint32_t ctr[4] = { (uint32_t)row, (uint32_t)(row >> 32), 0, 0x554E4946u };
	uint32_t w[4];
	vg_philox(ctr, key, w);
	return vg_u53(w[0], w[1]);
}

VG_GEN_FN int32_t vg_round_half_away(double x)
{
	double r = floor(fabs(x) + 0.5);
	return (int32_t)(x < 0 ? -r : r);
}

/* uniformi in [0,100): the high word of w0 * 100 */
VG_GEN_FN int32_t vg_gen_uniformi(uint32_t seed, uint64_t row)
{
	uint32_t key[2] = { seed, VG_GC_UNIFORMI };
	uint32_t ctr[4] = { (uint32_t)row, (uint32_t)(row >> 32), 0, 0x554E4946u };
	uint32_t w[4];
	vg_philox(ctr, key, w);
	return (int32_t)(((uint64_t)w[0] * 100u) >> 32);
}

/* uniformf in [0,100) as f32 (a product that rounds up to 100 is pulled
 * back to the largest f32 below it) */
VG_GEN_FN float vg_gen_uniformf(uint32_t seed, uint64_t row)
{
	float f = (float)(vg_gen_uniform(seed, VG_GC_UNIFORMF, row) * 100.0);
	return f < 100.0f ? f : 99.99999237060547f;
}

VG_GEN_FN int32_t vg_gen_normali(uint32_t seed, int column, double sigma, uint64_t row)
{
	return vg_round_half_away(vg_gen_normal(seed, (uint32_t)column, row) * sigma);
}

VG_GEN_FN float vg_gen_normalf(uint32_t seed, int column, double sigma, uint64_t row)
{
	return (float)(vg_gen_normal(seed, (uint32_t)column, row) * sigma);
}

#endif
