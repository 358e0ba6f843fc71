/*
 * Features of a (program, tablet) pair for the result-size predictor.
 *
 * For each tablet: the filter evaluated under independence, each
 * `column op constant` compare passing at its histogram estimate on the
 * tablet and any other compare at a value given by its kind (ordered, =,
 * !=) in each of seven variants: all at 1/2, 1/4 or 3/4, and the four
 * of include/virg_gpu.h (as logits scaled by 1/4, and the all-1/2
 * estimate itself); the program's shape
 * (compares estimated by histograms and not, keep and drop branches, Invalid
 * and logic ops, length); the tablet's rows; and the tablet's own column
 * statistics (min, max, 16-bin histogram per column, given).
 *
 * The walk follows the VM's control flow over the filter
 * (reference: src/vm/virginia.c:48-130, :424-440): a taken compare jumps to
 * p3 keeping validity when p4 is set and dropping the row otherwise; a row
 * that does not take it falls through; Invalid drops the rows reaching it;
 * jumps are forward only.  No counterpart in the reference otherwise.
 */
#include "internal.h"
#include "virg_gpu.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NU 7                       /* variants of the unknown compares' passing rates */
/* per variant, the rate of an ordered (or logic), an = and a != compare */
static const double U[NU][3] = {
	{ 0.5, 0.5, 0.5 }, { 0.25, 0.25, 0.25 }, { 0.75, 0.75, 0.75 },
	{ 0.5, 0.05, 0.95 }, { 0.25, 0.5, 0.5 }, { 0.75, 0.5, 0.5 }, { 0.5, 0.0, 1.0 },
};

enum { RD_EXPR = 0, RD_COL, RD_CONST };

/* fraction of each tablet's rows with column c < v (histogram, linear
 * inside a bin) */
static void cdf(const vg_table_stats *st, unsigned c, double v, double *out)
{
	const double *e = st->edges + (size_t)c * (VG_FEAT_BINS + 1);
	const unsigned nt = st->ntab;
	if (v <= e[0]) {
		for (unsigned t = 0; t < nt; t++)
			out[t] = 0.0;
		return;
	}
	if (v >= e[VG_FEAT_BINS]) {
		for (unsigned t = 0; t < nt; t++)
			out[t] = 1.0;
		return;
	}
	unsigned k = 0;                    /* last edge <= v */
	while (k + 1 <= VG_FEAT_BINS && e[k + 1] <= v)
		k++;
	if (k > VG_FEAT_BINS - 1)
		k = VG_FEAT_BINS - 1;
	const double part = (v - e[k]) / (e[k + 1] - e[k]);
	for (unsigned t = 0; t < nt; t++) {
		const double *h = st->hist + ((size_t)t * VG_FEAT_COLS + c) * VG_FEAT_BINS;
		double sum = 0.0;
		for (unsigned j = 0; j < k; j++)
			sum += h[j];
		out[t] = sum + part * h[k];
	}
}

/* estimated fraction of each tablet's rows with (column c op v), in [0, 1] */
static void passing(const vg_table_stats *st, unsigned c, int op, double v, double *out, double *tmp)
{
	const unsigned nt = st->ntab;
	double *lt = out, *le = tmp, *eq = tmp + nt;
	if (st->is_int[c]) {
		cdf(st, c, ceil(v), lt);
		cdf(st, c, floor(v) + 1.0, le);
		if (floor(v) == v) {
			cdf(st, c, floor(v), eq);
			for (unsigned t = 0; t < nt; t++)
				eq[t] = le[t] - eq[t];
		} else {
			for (unsigned t = 0; t < nt; t++)
				eq[t] = 0.0;
		}
	} else {
		cdf(st, c, v, lt);
		for (unsigned t = 0; t < nt; t++) {
			le[t] = lt[t];
			eq[t] = 0.0;
		}
	}
	for (unsigned t = 0; t < nt; t++) {
		double p;
		switch (op) {
		case OP_Lt: p = lt[t]; break;
		case OP_Le: p = le[t]; break;
		case OP_Gt: p = 1.0 - le[t]; break;
		case OP_Ge: p = 1.0 - lt[t]; break;
		case OP_Eq: p = eq[t]; break;
		default: p = 1.0 - eq[t]; break;
		}
		out[t] = p < 0.0 ? 0.0 : p > 1.0 ? 1.0 : p;
	}
}

static int is_cmp(int op)
{
	return op == OP_Lt || op == OP_Le || op == OP_Gt || op == OP_Ge || op == OP_Eq || op == OP_Neq;
}

static int mirror(int op)
{
	switch (op) {
	case OP_Lt: return OP_Gt;
	case OP_Le: return OP_Ge;
	case OP_Gt: return OP_Lt;
	case OP_Ge: return OP_Le;
	default: return op;
	}
}

static float scaled_logit(double p)
{
	const double q = p < 1e-4 ? 1e-4 : p > 1.0 - 1e-4 ? 1.0 - 1e-4 : p;
	return (float)(log(q / (1.0 - q)) / VG_FEAT_SKIP);
}

int vg_program_features(const virg_op *ops, unsigned num_ops, const vg_table_stats *st, float *out)
{
	const unsigned nt = st->ntab;
	VIRG_CHECK(nt == 0, "no tablets");
	unsigned start = 0, res = num_ops;
	for (unsigned i = 0; i < num_ops; i++)
		if (ops[i].op == OP_Parallel) {
			start = i + 1;
			break;
		}
	for (unsigned i = start; i < num_ops; i++)
		if (ops[i].op == OP_Result) {
			res = i;
			break;
		}
	/* arrive[pc][u][tablet]: fraction of rows reaching pc (have[pc]: any) */
	const size_t row = (size_t)NU * nt;
	double *arrive = (double *)calloc((size_t)(num_ops + 1) * row + 4 * (size_t)nt, sizeof(double));
	unsigned char *have = (unsigned char *)calloc(num_ops + 1, 1);
	if (!arrive || !have) {
		free(arrive);
		free(have);
		VIRG_CHECK(1, "Out of memory");
	}
	double *p = arrive + (size_t)(num_ops + 1) * row, *tmp = p + nt;
	int rkind[VIRG_REGS];
	double rval[VIRG_REGS];
	memset(rkind, 0, sizeof rkind);
	unsigned known = 0, unknown = 0, keep = 0, drop = 0, ninvalid = 0, nlogic = 0;
	for (size_t j = 0; j < row; j++)
		arrive[(size_t)start * row + j] = 1.0;
	have[start] = 1;
	for (unsigned pc = start; pc < res; pc++) {
		const virg_op *o = &ops[pc];
		const int k = o->op;
		const int m = have[pc];
		double *mp = arrive + (size_t)pc * row;
		int next = m;                      /* rows fall through to pc + 1 */
		const int r1 = o->p1 >= 0 && o->p1 < VIRG_REGS ? o->p1 : -1;
		if (k == OP_Column) {
			if (r1 >= 0) {
				rkind[r1] = RD_COL;
				rval[r1] = o->p2;
			}
		} else if (k == OP_Integer || k == OP_Float) {
			if (r1 >= 0) {
				rkind[r1] = RD_CONST;
				rval[r1] = k == OP_Float ? (double)o->p4.f : (double)o->p2;
			}
		} else if (k == OP_Invalid) {
			ninvalid++;
			next = 0;
		} else if (is_cmp(k) || k == OP_And || k == OP_Or || k == OP_Not) {
			int estimated = 0;
			if (is_cmp(k)) {
				const int r2 = o->p2 >= 0 && o->p2 < VIRG_REGS ? o->p2 : -1;
				int ak = r1 >= 0 ? rkind[r1] : RD_EXPR, bk = r2 >= 0 ? rkind[r2] : RD_EXPR;
				double av = r1 >= 0 ? rval[r1] : 0.0, bv = r2 >= 0 ? rval[r2] : 0.0;
				int op = k;
				if (ak == RD_CONST && bk == RD_COL) {
					int tk = ak; ak = bk; bk = tk;
					double tv = av; av = bv; bv = tv;
					op = mirror(k);
				}
				if (ak == RD_COL && av < VG_FEAT_COLS && bk == RD_CONST) {
					passing(st, (unsigned)av, op, bv, p, tmp);
					estimated = 1;
					known++;
				} else {
					unknown++;
				}
				if (o->p4.i)
					keep++;
				else
					drop++;
			} else {
				nlogic++;
			}
			const int kind = k == OP_Eq ? 1 : k == OP_Neq ? 2 : 0;
			if (m) {
				const int jumps = o->p4.i && o->p3 > (int)pc && o->p3 <= (int)num_ops;
				double *tp = jumps ? arrive + (size_t)o->p3 * row : NULL;
				for (unsigned u = 0; u < NU; u++)
					for (unsigned t = 0; t < nt; t++) {
						const double pr = estimated ? p[t] : U[u][kind];
						const double x = mp[u * nt + t];
						if (tp)
							tp[u * nt + t] += x * pr;
						mp[u * nt + t] = x * (1.0 - pr);
					}
				if (tp)
					have[o->p3] = 1;
			}
		} else if (k != OP_Table && r1 >= 0) {
			rkind[r1] = RD_EXPR;
		}
		if (next && m) {
			double *np = arrive + (size_t)(pc + 1) * row;
			for (size_t j = 0; j < row; j++)
				np[j] += mp[j];
			have[pc + 1] = 1;
		}
	}
	const double *est = have[res] ? arrive + (size_t)res * row : NULL;
	const float shape[7] = { known / 4.0f, unknown / 4.0f, keep / 4.0f, drop / 4.0f, (float)ninvalid,
		(float)nlogic, (float)(num_ops / 40.0) };
	for (unsigned t = 0; t < nt; t++) {
		float *f = out + (size_t)t * VG_FEATURES;
		double e[NU];
		for (unsigned u = 0; u < NU; u++) {
			const double x = est ? est[u * nt + t] : 0.0;
			e[u] = x < 0.0 ? 0.0 : x > 1.0 ? 1.0 : x;
		}
		f[0] = scaled_logit(e[0]);
		f[1] = scaled_logit(e[1]);
		f[2] = scaled_logit(e[2]);
		f[3] = (float)e[0];
		for (int j = 0; j < 7; j++)
			f[4 + j] = shape[j];
		f[11] = (float)(st->rows[t] / (double)(1 << 20));
		for (unsigned u = 3; u < NU; u++)
			f[12 + u - 3] = scaled_logit(e[u]);
		for (int j = 0; j < VG_FEAT_COLS * (VG_FEAT_BINS + 2); j++)
			f[16 + j] = (float)st->tabfeat[(size_t)t * VG_FEAT_COLS * (VG_FEAT_BINS + 2) + j];
	}
	free(arrive);
	free(have);
	return VIRG_SUCCESS;
}
