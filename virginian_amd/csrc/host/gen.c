/*
 * Deterministic synthetic tables.
 *
 * Schema (the reference's comparison table, db/generate.c:21-27, with the
 * distributions of the benchmark configuration): key id int64 = row index;
 * uniformi int32 in [0,100); normali5, normali20 int32 = round(N(0,1)*5),
 * round(N(0,1)*20) (half away from zero); uniformf f32 in [0,100);
 * normalf5, normalf20 f32 = N(0,1)*5, N(0,1)*20.
 *
 * Every value is a pure function of (seed, column, row): Philox4x32-10
 * (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
 * 1, 2, 3", SC'11) keyed by (seed, column) with counter (row, attempt),
 * and normals from the Marsaglia polar method with a logarithm written
 * here from IEEE add/multiply/divide only, so any machine (and numpy)
 * reproduces the tables bit for bit.  Compiled with -ffp-contract=off.
 */
#include "internal.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void virg_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
	uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
	uint32_t k0 = key[0], k1 = key[1];
	for (int i = 0; i < 10; i++) {
		uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
		uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
		uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
		uint32_t n1 = (uint32_t)p1;
		uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
		uint32_t n3 = (uint32_t)p0;
		c0 = n0; c1 = n1; c2 = n2; c3 = n3;
		k0 += PHILOX_W0;
		k1 += PHILOX_W1;
	}
	out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double u53(uint32_t a, uint32_t b)
{
	return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* natural logarithm of x in (0, 1]: x = m * 2^e with m in [sqrt(1/2),
 * sqrt(2)), log m = 2 atanh(f), f = (m-1)/(m+1), |f| < 0.1716, by its odd
 * series to f^19 (error below 1e-17 relative). */
double virg_log(double x)
{
	uint64_t bits;
	memcpy(&bits, &x, 8);
	int e = (int)((bits >> 52) & 0x7FF) - 1023;
	bits = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
	double m;
	memcpy(&m, &bits, 8);
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
double virg_gen_normal(uint32_t seed, uint32_t column, uint64_t row)
{
	uint32_t key[2] = { seed, column };
	for (uint32_t attempt = 0;; attempt++) {
		uint32_t ctr[4] = { (uint32_t)row, (uint32_t)(row >> 32), attempt, 0x4E4F524Du };
		uint32_t w[4];
		virg_philox(ctr, key, w);
		double a = 2.0 * u53(w[0], w[1]) - 1.0;
		double b = 2.0 * u53(w[2], w[3]) - 1.0;
		double s = a * a + b * b;
		if (s >= 1.0 || s == 0.0)
			continue;
		return a * sqrt(-2.0 * virg_log(s) / s);
	}
}

/* uniform double in [0,1) for (seed, column, row) */
double virg_gen_uniform(uint32_t seed, uint32_t column, uint64_t row)
{
	uint32_t key[2] = { seed, column };
	uint32_t ctr[4] = { (uint32_t)row, (uint32_t)(row >> 32), 0, 0x554E4946u };
	uint32_t w[4];
	virg_philox(ctr, key, w);
	return u53(w[0], w[1]);
}

static int32_t round_half_away(double x)
{
	double r = floor(fabs(x) + 0.5);
	return (int32_t)(x < 0 ? -r : r);
}

enum { GC_UNIFORMI = 1, GC_NORMALI5, GC_NORMALI20, GC_UNIFORMF, GC_NORMALF5, GC_NORMALF20 };

/* Fill rows [row0, row0+n) of the 7 standard columns. */
void virg_gen_rows(uint32_t seed, uint64_t row0, uint64_t n, int64_t *id, int32_t *ui,
	int32_t *ni5, int32_t *ni20, float *uf, float *nf5, float *nf20)
{
	for (uint64_t i = 0; i < n; i++) {
		uint64_t row = row0 + i;
		if (id) id[i] = (int64_t)row;
		if (ui) {
			uint32_t key[2] = { seed, GC_UNIFORMI };
			uint32_t ctr[4] = { (uint32_t)row, (uint32_t)(row >> 32), 0, 0x554E4946u };
			uint32_t w[4];
			virg_philox(ctr, key, w);
			ui[i] = (int32_t)(((uint64_t)w[0] * 100u) >> 32);
		}
		if (ni5) ni5[i] = round_half_away(virg_gen_normal(seed, GC_NORMALI5, row) * 5.0);
		if (ni20) ni20[i] = round_half_away(virg_gen_normal(seed, GC_NORMALI20, row) * 20.0);
		if (uf) {
			float f = (float)(virg_gen_uniform(seed, GC_UNIFORMF, row) * 100.0);
			uf[i] = f < 100.0f ? f : 99.99999237060547f;
		}
		if (nf5) nf5[i] = (float)(virg_gen_normal(seed, GC_NORMALF5, row) * 5.0);
		if (nf20) nf20[i] = (float)(virg_gen_normal(seed, GC_NORMALF20, row) * 20.0);
	}
}

typedef struct {
	uint32_t seed;
	uint64_t row0, n;
	int64_t *id;
	int32_t *ui, *ni5, *ni20;
	float *uf, *nf5, *nf20;
} genjob;

static void *gen_worker(void *arg)
{
	genjob *j = (genjob *)arg;
	virg_gen_rows(j->seed, j->row0, j->n, j->id, j->ui, j->ni5, j->ni20, j->uf, j->nf5, j->nf20);
	return NULL;
}

/* Parallel version over nthreads host threads. */
void virg_gen_rows_mt(uint32_t seed, uint64_t row0, uint64_t n, int64_t *id, int32_t *ui,
	int32_t *ni5, int32_t *ni20, float *uf, float *nf5, float *nf20, int nthreads)
{
	if (nthreads < 1)
		nthreads = 1;
	if (nthreads > 256)
		nthreads = 256;
	pthread_t th[256];
	int live[256];
	genjob jobs[256];
	uint64_t per = (n + (uint64_t)nthreads - 1) / (uint64_t)nthreads;
	for (int t = 0; t < nthreads; t++) {
		live[t] = 0;
		uint64_t b = (uint64_t)t * per;
		if (b >= n)
			continue;
		uint64_t c = n - b < per ? n - b : per;
		genjob *j = &jobs[t];
		j->seed = seed;
		j->row0 = row0 + b;
		j->n = c;
		j->id = id ? id + b : NULL;
		j->ui = ui ? ui + b : NULL;
		j->ni5 = ni5 ? ni5 + b : NULL;
		j->ni20 = ni20 ? ni20 + b : NULL;
		j->uf = uf ? uf + b : NULL;
		j->nf5 = nf5 ? nf5 + b : NULL;
		j->nf20 = nf20 ? nf20 + b : NULL;
		if (pthread_create(&th[t], NULL, gen_worker, j) != 0)
			gen_worker(j);
		else
			live[t] = 1;
	}
	for (int t = 0; t < nthreads; t++)
		if (live[t])
			pthread_join(th[t], NULL);
}

/* Create table `name` with the standard schema and `rows` rows in tablets
 * of `tablet_rows` rows (0 = reference 8 MB tablets filled through the
 * row-at-a-time insert path). */
int virg_gen_table(virginian *v, const char *name, uint64_t rows, unsigned tablet_rows,
	uint32_t seed, int nthreads)
{
	VIRG_CHECK(virg_table_create(v, name, VIRG_INT64) == VIRG_FAIL, "table create failed");
	unsigned tid;
	VIRG_CHECK(virg_table_getid(v, name, &tid) == VIRG_FAIL, "table lookup failed");
	static const char *names[6] = { "uniformi", "normali5", "normali20", "uniformf", "normalf5", "normalf20" };
	static const virg_t types[6] = { VIRG_INT, VIRG_INT, VIRG_INT, VIRG_FLOAT, VIRG_FLOAT, VIRG_FLOAT };
	for (int c = 0; c < 6; c++)
		VIRG_CHECK(virg_table_addcolumn(v, tid, names[c], types[c]) == VIRG_FAIL, "add column failed");
	uint64_t chunk = tablet_rows ? tablet_rows : 65536;
	int64_t *id = (int64_t *)malloc(chunk * 8);
	int32_t *ui = (int32_t *)malloc(chunk * 4), *ni5 = (int32_t *)malloc(chunk * 4), *ni20 = (int32_t *)malloc(chunk * 4);
	float *uf = (float *)malloc(chunk * 4), *nf5 = (float *)malloc(chunk * 4), *nf20 = (float *)malloc(chunk * 4);
	int ok = id && ui && ni5 && ni20 && uf && nf5 && nf20;
	/* VIRG_GEN_PROGRESS=n: a line on stderr every n rows (long tables) */
	const char *pe = getenv("VIRG_GEN_PROGRESS");
	const uint64_t every = pe ? strtoull(pe, NULL, 10) : 0;
	for (uint64_t r0 = 0; ok && r0 < rows; r0 += chunk) {
		uint64_t n = rows - r0 < chunk ? rows - r0 : chunk;
		virg_gen_rows_mt(seed, r0, n, id, ui, ni5, ni20, uf, nf5, nf20, nthreads);
		if (every && (r0 + n) / every != r0 / every) {
			fprintf(stderr, "# generated %llu of %llu rows\n", (unsigned long long)(r0 + n),
				(unsigned long long)rows);
			fflush(stderr);
		}
		if (tablet_rows) {
			const void *cols[6] = { ui, ni5, ni20, uf, nf5, nf20 };
			unsigned cap = (tablet_rows + 15u) & ~15u;
			ok = virg_table_appendtablet(v, tid, cap, (unsigned)n, id, cols) == VIRG_SUCCESS;
		} else {
			for (uint64_t i = 0; i < n && ok; i++) {
				char data[24];
				memcpy(data + 0, &ui[i], 4);
				memcpy(data + 4, &ni5[i], 4);
				memcpy(data + 8, &ni20[i], 4);
				memcpy(data + 12, &uf[i], 4);
				memcpy(data + 16, &nf5[i], 4);
				memcpy(data + 20, &nf20[i], 4);
				ok = virg_table_insert(v, tid, (char *)&id[i], data, NULL) == VIRG_SUCCESS;
			}
		}
	}
	free(id); free(ui); free(ni5); free(ni20); free(uf); free(nf5); free(nf20);
	VIRG_CHECK(!ok, "table generation failed");
	return VIRG_SUCCESS;
}

/* Up to `max` tablet pointers of a table, in chain order (the tablets stay
 * resident; no lock is kept).  Returns the number of tablets. */
int virg_table_tablets(virginian *v, unsigned table_id, const virg_tablet_meta **out, int max)
{
	virg_tablet_meta *t;
	if (virg_db_load(v, v->db.first_tablet[table_id], &t) == VIRG_FAIL)
		return 0;
	int n = 0;
	while (1) {
		if (n < max)
			out[n] = t;
		n++;
		if (t->last_tablet)
			break;
		virg_db_loadnext(v, &t);
	}
	virg_tablet_unlock(v, t->id);
	return n;
}
