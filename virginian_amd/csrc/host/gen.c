/*
 * Deterministic synthetic tables.
 *
 * Schema (the reference's comparison table, db/generate.c:21-27, with the
 * distributions of the benchmark configuration): key id int64 = row index;
 * uniformi int32 in [0,100); normali5, normali20 int32 = round(N(0,1)*5),
 * round(N(0,1)*20) (half away from zero); uniformf f32 in [0,100);
 * normalf5, normalf20 f32 = N(0,1)*5, N(0,1)*20.  The values come from
 * gen_math.h, which the device table builder (gpu/gen.hip) compiles too.
 * Compiled with -ffp-contract=off.
 */
#include "internal.h"
#include "gen_math.h"
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

void virg_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
	vg_philox(ctr, key, out);
}

double virg_log(double x)
{
	return vg_log(x);
}

double virg_gen_normal(uint32_t seed, uint32_t column, uint64_t row)
{
	return vg_gen_normal(seed, column, row);
}

double virg_gen_uniform(uint32_t seed, uint32_t column, uint64_t row)
{
	return vg_gen_uniform(seed, column, row);
}

/* Fill rows [row0, row0+n) of the 7 standard columns. */
void virg_gen_rows(uint32_t seed, uint64_t row0, uint64_t n, int64_t *id, int32_t *ui,
	int32_t *ni5, int32_t *ni20, float *uf, float *nf5, float *nf20)
{
	for (uint64_t i = 0; i < n; i++) {
		uint64_t row = row0 + i;
		if (id) id[i] = (int64_t)row;
		if (ui) ui[i] = vg_gen_uniformi(seed, row);
		if (ni5) ni5[i] = vg_gen_normali(seed, VG_GC_NORMALI5, 5.0, row);
		if (ni20) ni20[i] = vg_gen_normali(seed, VG_GC_NORMALI20, 20.0, row);
		if (uf) uf[i] = vg_gen_uniformf(seed, row);
		if (nf5) nf5[i] = vg_gen_normalf(seed, VG_GC_NORMALF5, 5.0, row);
		if (nf20) nf20[i] = vg_gen_normalf(seed, VG_GC_NORMALF20, 20.0, row);
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
/* Create table `name` with the standard schema (no rows yet). */
int virg_gen_schema(virginian *v, const char *name, unsigned *table_id)
{
	VIRG_CHECK(virg_table_create(v, name, VIRG_INT64) == VIRG_FAIL, "table create failed");
	unsigned tid;
	VIRG_CHECK(virg_table_getid(v, name, &tid) == VIRG_FAIL, "table lookup failed");
	static const char *names[6] = { "uniformi", "normali5", "normali20", "uniformf", "normalf5", "normalf20" };
	static const virg_t types[6] = { VIRG_INT, VIRG_INT, VIRG_INT, VIRG_FLOAT, VIRG_FLOAT, VIRG_FLOAT };
	for (int c = 0; c < 6; c++)
		VIRG_CHECK(virg_table_addcolumn(v, tid, names[c], types[c]) == VIRG_FAIL, "add column failed");
	table_id[0] = tid;
	return VIRG_SUCCESS;
}

int virg_gen_table(virginian *v, const char *name, uint64_t rows, unsigned tablet_rows,
	uint32_t seed, int nthreads)
{
	unsigned tid;
	VIRG_CHECK(virg_gen_schema(v, name, &tid) == VIRG_FAIL, "table create failed");
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
