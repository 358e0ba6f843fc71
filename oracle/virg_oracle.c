/*
 * ORACLE -- test infrastructure only.  Nothing in the product library links
 * or calls this file; tests, the smoke check and the benchmark's CPU
 * baseline use it as the checker.
 *
 * A C restatement of the reference CPU VM's data-parallel interpreter
 * (reference: src/vm/virginia.c:228-891, driven per tablet as in
 * src/vm/vm_cpu.c:31-94).  It keeps the reference's execution model:
 * 64-row blocks; a per-row program counter and validity flag; Column and
 * Rowid copy the whole block (virginia.c:442-474); Integer, Float and
 * arithmetic write rows whose pc is the current op (:386-412, :138-194);
 * compares move a row whose compare is true to p3 and set its validity to
 * p4 if it was valid (:48-130); Result emits valid rows in row order and
 * advances every row (:476-578); Converge ends the block (:381-384).
 *
 * Additions for checking: each block tracks, per register, which rows have
 * had it written in that block; a row that reads a register it never wrote
 * (and so would see a value left over from another row or block) marks
 * the program "ill-formed" for parity purposes.  Assertion failures of the
 * reference (operand type mismatch, :52 and :143; a row pc behind the
 * block pc, :376) are returned as errors instead of aborting.  Cast uses
 * the ops own p1/p2 (the reference's CPU Cast reads stale p1/p2, :682).
 *
 * Two drivers: oracle_select (single thread, tablets in chain order: the
 * reference single-core order) and oracle_select_mt (whole tablets handed
 * to threads greedily, results reserved per block under a mutex, the
 * reference's multi-core scheme, :300-349 and :483-512; row order across
 * blocks is not deterministic, the row multiset is).  oracle_select_mt
 * with ordered=1 writes each tablet's rows into its own range so the
 * output equals the single-thread order.
 */
#include "virg_oracle.h"
#include <pthread.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#define SIMD 64

typedef union {
	int i[SIMD];
	long long li[SIMD];
	float f[SIMD];
	double d[SIMD];
	char c[SIMD];
} regblock;

typedef struct {
	regblock reg[VIRG_REGS];
	virg_t type[VIRG_REGS];
	size_t stride[VIRG_REGS];
	uint64_t written[VIRG_REGS];   /* rows that wrote the register in this block */
	unsigned row_pc[SIMD];
	int valid[SIMD];
} ctx;

typedef struct {
	/* output */
	void **out;
	unsigned nout;
	unsigned long long cap;
	unsigned long long rows;
	/* status */
	int illformed;
	int failed;
	char err[160];
	pthread_mutex_t *lock;        /* multi-threaded reservation */
} sink;

static void fail(sink *s, const char *fmt, ...)
{
	if (s->failed)
		return;
	s->failed = 1;
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(s->err, sizeof s->err, fmt, ap);
	va_end(ap);
}

static int tsize(int t)
{
	switch (t) {
	case VIRG_INT: case VIRG_FLOAT: return 4;
	case VIRG_INT64: case VIRG_DOUBLE: return 8;
	case VIRG_CHAR: return 1;
	}
	return 0;
}

#define ACTIVE(i) (C->row_pc[i] == pc)

/* compare (or and/or) of registers p1, p2 typed by p1 */
#define REGCMP(OP)                                                                   \
	do {                                                                             \
		int p1 = o->p1, p2 = o->p2, p3 = o->p3;                                      \
		if (C->type[p1] != C->type[p2]) { fail(s, "op %u: operand types differ", pc); return; } \
		for (unsigned i = 0; i < n; i++) {                                           \
			if (!ACTIVE(i)) continue;                                                \
			if (!((C->written[p1] >> i) & 1) || !((C->written[p2] >> i) & 1))        \
				s->illformed = 1;                                                    \
			int x = 0;                                                               \
			switch (C->type[p1]) {                                                   \
			case VIRG_INT: x = C->reg[p1].i[i] OP C->reg[p2].i[i]; break;            \
			case VIRG_FLOAT: x = C->reg[p1].f[i] OP C->reg[p2].f[i]; break;          \
			case VIRG_INT64: x = C->reg[p1].li[i] OP C->reg[p2].li[i]; break;        \
			case VIRG_DOUBLE: x = C->reg[p1].d[i] OP C->reg[p2].d[i]; break;         \
			case VIRG_CHAR: x = C->reg[p1].c[i] OP C->reg[p2].c[i]; break;           \
			default: fail(s, "op %u: bad type", pc); return;                         \
			}                                                                        \
			if (x) {                                                                 \
				if (C->valid[i]) C->valid[i] = o->p4.i;                              \
				C->row_pc[i] = (unsigned)p3;                                         \
			} else {                                                                 \
				C->row_pc[i]++;                                                      \
			}                                                                        \
		}                                                                            \
	} while (0)

#define MATHOP(OP)                                                                   \
	do {                                                                             \
		int p1 = o->p1, p2 = o->p2, p3 = o->p3;                                      \
		if (C->type[p2] != C->type[p3]) { fail(s, "op %u: operand types differ", pc); return; } \
		C->type[p1] = C->type[p2];                                                   \
		C->stride[p1] = C->stride[p2];                                               \
		for (unsigned i = 0; i < n; i++) {                                           \
			if (!ACTIVE(i)) continue;                                                \
			if (!((C->written[p2] >> i) & 1) || !((C->written[p3] >> i) & 1))        \
				s->illformed = 1;                                                    \
			switch (C->type[p1]) {                                                   \
			case VIRG_INT: C->reg[p1].i[i] = (int)((unsigned)C->reg[p2].i[i] OP (unsigned)C->reg[p3].i[i]); break; \
			case VIRG_FLOAT: C->reg[p1].f[i] = C->reg[p2].f[i] OP C->reg[p3].f[i]; break; \
			case VIRG_INT64: C->reg[p1].li[i] = (long long)((unsigned long long)C->reg[p2].li[i] OP (unsigned long long)C->reg[p3].li[i]); break; \
			case VIRG_DOUBLE: C->reg[p1].d[i] = C->reg[p2].d[i] OP C->reg[p3].d[i]; break; \
			case VIRG_CHAR: C->reg[p1].c[i] = (char)(C->reg[p2].c[i] OP C->reg[p3].c[i]); break; \
			default: fail(s, "op %u: bad type", pc); return;                         \
			}                                                                        \
			C->row_pc[i]++;                                                          \
			C->written[p1] |= 1ull << i;                                             \
		}                                                                            \
	} while (0)

static void divide(ctx *C, sink *s, const virg_op *o, unsigned pc, unsigned n)
{
	int p1 = o->p1, p2 = o->p2, p3 = o->p3;
	if (C->type[p2] != C->type[p3]) {
		fail(s, "op %u: operand types differ", pc);
		return;
	}
	C->type[p1] = C->type[p2];
	C->stride[p1] = C->stride[p2];
	for (unsigned i = 0; i < n; i++) {
		if (!ACTIVE(i))
			continue;
		if (!((C->written[p2] >> i) & 1) || !((C->written[p3] >> i) & 1))
			s->illformed = 1;
		switch (C->type[p1]) {
		case VIRG_INT:
			if (C->reg[p3].i[i] == 0 || (C->reg[p3].i[i] == -1 && C->reg[p2].i[i] == (-2147483647 - 1))) {
				fail(s, "op %u: integer division fault", pc);
				return;
			}
			C->reg[p1].i[i] = C->reg[p2].i[i] / C->reg[p3].i[i];
			break;
		case VIRG_FLOAT: C->reg[p1].f[i] = C->reg[p2].f[i] / C->reg[p3].f[i]; break;
		case VIRG_INT64:
			if (C->reg[p3].li[i] == 0 || (C->reg[p3].li[i] == -1 && C->reg[p2].li[i] == (-9223372036854775807LL - 1))) {
				fail(s, "op %u: integer division fault", pc);
				return;
			}
			C->reg[p1].li[i] = C->reg[p2].li[i] / C->reg[p3].li[i];
			break;
		case VIRG_DOUBLE: C->reg[p1].d[i] = C->reg[p2].d[i] / C->reg[p3].d[i]; break;
		case VIRG_CHAR:
			if (C->reg[p3].c[i] == 0) {
				fail(s, "op %u: integer division fault", pc);
				return;
			}
			C->reg[p1].c[i] = (char)(C->reg[p2].c[i] / C->reg[p3].c[i]);
			break;
		default: fail(s, "op %u: bad type", pc); return;
		}
		C->row_pc[i]++;
		C->written[p1] |= 1ull << i;
	}
}

static double rd(const ctx *C, int r, unsigned i, long long *li)
{
	switch (C->type[r]) {
	case VIRG_INT: *li = C->reg[r].i[i]; return (double)C->reg[r].i[i];
	case VIRG_INT64: *li = C->reg[r].li[i]; return (double)C->reg[r].li[i];
	case VIRG_FLOAT: *li = (long long)C->reg[r].f[i]; return (double)C->reg[r].f[i];
	case VIRG_DOUBLE: *li = (long long)C->reg[r].d[i]; return C->reg[r].d[i];
	case VIRG_CHAR: *li = C->reg[r].c[i]; return (double)C->reg[r].c[i];
	}
	*li = 0;
	return 0;
}

static void cast(ctx *C, sink *s, const virg_op *o, unsigned pc, unsigned n)
{
	/* Cast p2 to type p1 in place, row by row, for the rows at this op.
	 * The register block is a union of per-type arrays, so the converted
	 * block is built apart and copied back: converting in place in row
	 * order would let a widened row overwrite rows not yet read.  (The
	 * reference's Cast, virginia.c:682-862, reads p1/p2 left over from the
	 * previous op and is never generated; this is the clean definition.) */
	int to = o->p1, r = o->p2;
	int from = C->type[r];
	regblock out = C->reg[r];
	const int resized = tsize(to) != tsize(from);
	if (resized)
		memset(&out, 0, sizeof out);
	for (unsigned i = 0; i < n; i++) {
		if (!ACTIVE(i)) {
			/* a row that skips a cast changing the width has no value of
			 * the new type: reading it later is ill-formed */
			if (resized)
				C->written[r] &= ~(1ull << i);
			continue;
		}
		if (!((C->written[r] >> i) & 1))
			s->illformed = 1;
		long long li;
		double d = rd(C, r, i, &li);
		int isint = from == VIRG_INT || from == VIRG_INT64 || from == VIRG_CHAR;
		float f = from == VIRG_FLOAT ? C->reg[r].f[i] : (from == VIRG_DOUBLE ? (float)C->reg[r].d[i] : (float)li);
		switch (to) {
		case VIRG_INT: out.i[i] = isint ? (int)li : (from == VIRG_FLOAT ? (int)C->reg[r].f[i] : (int)d); break;
		case VIRG_INT64: out.li[i] = isint ? li : (from == VIRG_FLOAT ? (long long)C->reg[r].f[i] : (long long)d); break;
		case VIRG_FLOAT: out.f[i] = f; break;
		case VIRG_DOUBLE: out.d[i] = isint ? (double)li : d; break;
		case VIRG_CHAR: out.c[i] = isint ? (char)li : (char)(int)d; break;
		default: fail(s, "op %u: bad cast", pc); return;
		}
		C->row_pc[i]++;
		C->written[r] |= 1ull << i;
	}
	C->reg[r] = out;
	C->type[r] = (virg_t)to;
	C->stride[r] = (size_t)tsize(to);
}

/* Emit the valid rows of a block.  The multi-threaded driver reserves its
 * range under the mutex (reference virginia.c:488-512). */
static void result(ctx *C, sink *s, const virg_op *o, unsigned pc, unsigned n, unsigned long long *wr)
{
	unsigned total = 0;
	for (unsigned i = 0; i < n; i++)
		if (C->valid[i])
			total++;
	unsigned long long w;
	if (wr) {
		w = *wr;
		*wr += total;
	} else if (s->lock) {
		pthread_mutex_lock(s->lock);
		w = s->rows;
		s->rows += total;
		pthread_mutex_unlock(s->lock);
	} else {
		w = s->rows;
		s->rows += total;
	}
	int p1 = o->p1, p2 = o->p2;
	if (p1 < 0 || p2 < 0 || p1 + p2 > VIRG_REGS) {
		fail(s, "op %u: Result registers out of range", pc);
		return;
	}
	for (int j = p1; j < p1 + p2; j++) {
		size_t st = C->stride[j];
		unsigned long long row = w;
		for (unsigned i = 0; i < n; i++) {
			if (!C->valid[i])
				continue;
			if (!((C->written[j] >> i) & 1))
				s->illformed = 1;
			if (s->out && (unsigned)(j - p1) < s->nout && s->out[j - p1] && row < s->cap)
				memcpy((char *)s->out[j - p1] + row * st, (char *)&C->reg[j] + st * i, st);
			row++;
		}
	}
	if (w + total > s->cap && s->out)
		fail(s, "result larger than the output capacity");
	for (unsigned i = 0; i < n; i++)
		C->row_pc[i]++;
}

/* Run the data-parallel section over one tablet. */
static void run_tablet(const virg_op *stmt, unsigned num_ops, unsigned pc0,
	const virg_tablet_meta *tab, sink *s, unsigned long long *wr)
{
	ctx *C = (ctx *)calloc(1, sizeof(ctx));
	if (!C) {
		fail(s, "out of memory");
		return;
	}
	for (unsigned row = 0; row < tab->rows && !s->failed; row += SIMD) {
		unsigned n = tab->rows - row < SIMD ? tab->rows - row : SIMD;
		for (unsigned i = 0; i < SIMD; i++) {
			C->row_pc[i] = pc0;
			C->valid[i] = i < n;
		}
		for (int r = 0; r < VIRG_REGS; r++)
			C->written[r] = 0;
		const uint64_t all = n == 64 ? ~0ull : ((1ull << n) - 1);
		unsigned pc = pc0;
		while (!s->failed) {
			if (pc >= num_ops) {
				fail(s, "ran past the end of the program");
				break;
			}
			for (unsigned i = 0; i < n; i++)
				if (C->row_pc[i] < pc) {
					fail(s, "op %u: row behind the block pc (backward jump)", pc);
					break;
				}
			const virg_op *o = &stmt[pc];
			if (o->op == OP_Converge)
				break;
			switch (o->op) {
			case OP_Integer:
			case OP_Float:
				for (unsigned i = 0; i < n; i++)
					if (ACTIVE(i)) {
						if (o->op == OP_Integer)
							C->reg[o->p1].i[i] = o->p2;
						else
							C->reg[o->p1].f[i] = o->p4.f;
						C->row_pc[i]++;
						C->written[o->p1] |= 1ull << i;
					}
				C->type[o->p1] = o->op == OP_Integer ? VIRG_INT : VIRG_FLOAT;
				C->stride[o->p1] = 4;
				break;
			case OP_Invalid:
				for (unsigned i = 0; i < n; i++)
					if (ACTIVE(i)) {
						C->valid[i] = 0;
						C->row_pc[i]++;
					}
				break;
			case OP_Le: REGCMP(<=); break;
			case OP_Lt: REGCMP(<); break;
			case OP_Ge: REGCMP(>=); break;
			case OP_Gt: REGCMP(>); break;
			case OP_Eq: REGCMP(==); break;
			case OP_Neq: REGCMP(!=); break;
			case OP_And: REGCMP(&&); break;
			case OP_Or: REGCMP(||); break;
			case OP_Column: {
				if (o->p2 < 0 || (unsigned)o->p2 >= tab->fixed_columns) {
					fail(s, "op %u: no column %d", pc, o->p2);
					break;
				}
				size_t st = tab->fixed_stride[o->p2];
				memcpy(&C->reg[o->p1], (const char *)tab + tab->fixed_block + tab->fixed_offset[o->p2] + st * row, st * n);
				C->type[o->p1] = tab->fixed_type[o->p2];
				C->stride[o->p1] = st;
				C->written[o->p1] = all;
				for (unsigned i = 0; i < n; i++)
					if (ACTIVE(i))
						C->row_pc[i]++;
				break;
			}
			case OP_Rowid: {
				size_t st = tab->key_stride;
				memcpy(&C->reg[o->p1], (const char *)tab + tab->key_block + st * row, st * n);
				C->type[o->p1] = tab->key_type;
				C->stride[o->p1] = st;
				C->written[o->p1] = all;
				for (unsigned i = 0; i < n; i++)
					if (ACTIVE(i))
						C->row_pc[i]++;
				break;
			}
			case OP_Result:
				result(C, s, o, pc, n, wr);
				break;
			case OP_Add: MATHOP(+); break;
			case OP_Sub: MATHOP(-); break;
			case OP_Mul: MATHOP(*); break;
			case OP_Div: divide(C, s, o, pc, n); break;
			case OP_Not:
				for (unsigned i = 0; i < n; i++) {
					if (!ACTIVE(i))
						continue;
					if (!((C->written[o->p1] >> i) & 1))
						s->illformed = 1;
					int z = 0;
					switch (C->type[o->p1]) {
					case VIRG_INT: z = !C->reg[o->p1].i[i]; break;
					case VIRG_FLOAT: z = !C->reg[o->p1].f[i]; break;
					case VIRG_INT64: z = !C->reg[o->p1].li[i]; break;
					case VIRG_DOUBLE: z = !C->reg[o->p1].d[i]; break;
					case VIRG_CHAR: z = !C->reg[o->p1].c[i]; break;
					}
					if (z) {
						if (C->valid[i])
							C->valid[i] = o->p4.i;
						C->row_pc[i] = (unsigned)o->p3;
					} else {
						C->row_pc[i]++;
					}
				}
				break;
			case OP_Cast:
				cast(C, s, o, pc, n);
				break;
			default:
				fail(s, "op %u: invalid op %d in the data-parallel section", pc, o->op);
				break;
			}
			pc++;
		}
	}
	free(C);
}

long long oracle_select(const virg_op *stmt, unsigned num_ops, unsigned pc,
	const virg_tablet_meta *const *tabs, unsigned ntab, void **out, unsigned nout,
	unsigned long long cap, int *illformed, char *err, size_t errlen)
{
	sink s;
	memset(&s, 0, sizeof s);
	s.out = out;
	s.nout = nout;
	s.cap = cap;
	for (unsigned t = 0; t < ntab && !s.failed; t++)
		run_tablet(stmt, num_ops, pc, tabs[t], &s, NULL);
	if (illformed)
		*illformed = s.illformed;
	if (s.failed) {
		if (err && errlen)
			snprintf(err, errlen, "%s", s.err);
		return -1;
	}
	return (long long)s.rows;
}

typedef struct {
	const virg_op *stmt;
	unsigned num_ops, pc;
	const virg_tablet_meta *const *tabs;
	unsigned ntab;
	unsigned next;
	pthread_mutex_t tab_lock;
	sink *s;
	const unsigned long long *starts;   /* ordered mode: first output row per tablet */
} mtjob;

static void *mt_worker(void *arg)
{
	mtjob *j = (mtjob *)arg;
	while (1) {
		pthread_mutex_lock(&j->tab_lock);
		unsigned t = j->next++;
		pthread_mutex_unlock(&j->tab_lock);
		if (t >= j->ntab || j->s->failed)
			break;
		if (j->starts) {
			unsigned long long w = j->starts[t];
			run_tablet(j->stmt, j->num_ops, j->pc, j->tabs[t], j->s, &w);
		} else {
			run_tablet(j->stmt, j->num_ops, j->pc, j->tabs[t], j->s, NULL);
		}
	}
	return NULL;
}

long long oracle_select_mt(const virg_op *stmt, unsigned num_ops, unsigned pc,
	const virg_tablet_meta *const *tabs, unsigned ntab, void **out, unsigned nout,
	unsigned long long cap, int nthreads, int ordered, int *illformed, char *err, size_t errlen)
{
	sink s;
	memset(&s, 0, sizeof s);
	pthread_mutex_t res_lock;
	pthread_mutex_init(&res_lock, NULL);
	s.out = out;
	s.nout = nout;
	s.cap = cap;
	s.lock = &res_lock;
	unsigned long long *starts = NULL;
	if (ordered) {
		/* a counting pass gives every tablet its output range */
		starts = (unsigned long long *)calloc(ntab + 1, sizeof *starts);
		if (!starts)
			return -1;
		unsigned long long acc = 0;
		for (unsigned t = 0; t < ntab; t++) {
			starts[t] = acc;
			char e2[8];
			long long c = oracle_select(stmt, num_ops, pc, &tabs[t], 1, NULL, 0, ~0ull, NULL, e2, sizeof e2);
			if (c < 0) {
				free(starts);
				return oracle_select(stmt, num_ops, pc, tabs, ntab, out, nout, cap, illformed, err, errlen);
			}
			acc += (unsigned long long)c;
		}
		s.rows = acc;
	}
	if (nthreads < 1)
		nthreads = 1;
	if (nthreads > 512)
		nthreads = 512;
	mtjob j;
	memset(&j, 0, sizeof j);
	j.stmt = stmt;
	j.num_ops = num_ops;
	j.pc = pc;
	j.tabs = tabs;
	j.ntab = ntab;
	j.s = &s;
	j.starts = starts;
	pthread_mutex_init(&j.tab_lock, NULL);
	pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
	int *live = (int *)calloc((size_t)nthreads, sizeof(int));
	for (int i = 0; th && live && i < nthreads; i++)
		live[i] = pthread_create(&th[i], NULL, mt_worker, &j) == 0;
	if (th && live)
		for (int i = 0; i < nthreads; i++)
			if (live[i])
				pthread_join(th[i], NULL);
	free(th);
	free(live);
	free(starts);
	pthread_mutex_destroy(&j.tab_lock);
	pthread_mutex_destroy(&res_lock);
	if (illformed)
		*illformed = s.illformed;
	if (s.failed) {
		if (err && errlen)
			snprintf(err, errlen, "%s", s.err);
		return -1;
	}
	return (long long)s.rows;
}
