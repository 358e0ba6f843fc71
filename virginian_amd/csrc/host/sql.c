/*
 * SQL front-end: lexer, recursive-descent parser and op-program generator.
 *
 * Language (reference: src/sql/sql.l, src/sql/sql.y:78-200):
 *   SELECT resultcol {, resultcol} FROM table [WHERE conditions]
 *   resultcol  := expr [[AS] label]
 *   conditions := condition [(AND|OR) conditions]
 *              |  '(' condition (AND|OR) conditions ')'
 *   condition  := expr cmp expr | '(' expr cmp expr ')'
 *   expr       := '(' expr ')' | expr (+|-|*|/) expr | int | float | column
 * The reference grammar gives the arithmetic rule no precedence (its four
 * %expect'ed shift/reduce conflicts resolve to shift), so arithmetic is
 * right-associative with equal precedence: a*b+c means a*(b+c).  Condition
 * chains fold to the right: A and B or C means A and (B or C).  A
 * parenthesised condition group may only end a chain.  Integer literals
 * carry an optional leading '-', so "x -5" does not parse.
 *
 * Code generation restates src/sql/generate.c: pass 0 types, pass 1
 * constant folding (mixed int/float folds through an int conversion of both
 * operands, generate.c:334-343), pass 2 structure with common
 * subexpressions sharing a register (generate.c:394-464), conditions as a
 * forward jump chain ending in Invalid (generate.c:477-578,627-641), result
 * registers renumbered to be contiguous at the end (generate.c:650-674),
 * pass 3/4 index and jump resolution, pass 5 emit.
 */
#include "internal.h"
#include <ctype.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ lexer */
enum {
	T_END, T_SELECT, T_FROM, T_WHERE, T_AS, T_AND, T_OR, T_INT, T_FLOAT,
	T_NAME, T_PLUS, T_MINUS, T_MUL, T_DIV, T_EQ, T_NE, T_LT, T_LE, T_GT,
	T_GE, T_COMMA, T_LP, T_RP
};

typedef struct {
	int type;
	int ival;
	float fval;
	char name[64];
} token;

#define MAX_TOKENS 512

static int kw(const char *s, size_t n, const char *word)
{
	if (strlen(word) != n)
		return 0;
	for (size_t i = 0; i < n; i++)
		if (tolower((unsigned char)s[i]) != word[i])
			return 0;
	return 1;
}

/* Longest match, keywords before names on equal length (flex rules). */
static int lex(const char *q, token *toks, int *ntok, char *err, size_t errlen)
{
	int n = 0;
	const char *p = q;
	while (*p) {
		if (n >= MAX_TOKENS - 1) {
			snprintf(err, errlen, "query too long");
			return 0;
		}
		token *t = &toks[n];
		memset(t, 0, sizeof(*t));
		char c = *p;
		if (c == ' ' || c == '\t' || c == '\n') {
			p++;
			continue;
		}
		if (isalpha((unsigned char)c)) {
			const char *s = p;
			while (isalnum((unsigned char)*p))
				p++;
			size_t len = (size_t)(p - s);
			if (kw(s, len, "select")) t->type = T_SELECT;
			else if (kw(s, len, "from")) t->type = T_FROM;
			else if (kw(s, len, "where")) t->type = T_WHERE;
			else if (kw(s, len, "as")) t->type = T_AS;
			else if (kw(s, len, "and")) t->type = T_AND;
			else if (kw(s, len, "or")) t->type = T_OR;
			else {
				t->type = T_NAME;
				if (len >= sizeof(t->name))
					len = sizeof(t->name) - 1;
				memcpy(t->name, s, len);
				t->name[len] = '\0';
			}
			n++;
			continue;
		}
		if (isdigit((unsigned char)c) || (c == '-' && isdigit((unsigned char)p[1]))) {
			const char *s = p;
			if (*p == '-')
				p++;
			while (isdigit((unsigned char)*p))
				p++;
			if (*p == '.' && isdigit((unsigned char)p[1])) {
				p++;
				while (isdigit((unsigned char)*p))
					p++;
				char buf[96];
				size_t len = (size_t)(p - s) < sizeof(buf) - 1 ? (size_t)(p - s) : sizeof(buf) - 1;
				memcpy(buf, s, len);
				buf[len] = '\0';
				t->type = T_FLOAT;
				t->fval = (float)atof(buf);
			} else {
				char buf[96];
				size_t len = (size_t)(p - s) < sizeof(buf) - 1 ? (size_t)(p - s) : sizeof(buf) - 1;
				memcpy(buf, s, len);
				buf[len] = '\0';
				t->type = T_INT;
				t->ival = atoi(buf);
			}
			n++;
			continue;
		}
		p++;
		switch (c) {
		case '=': t->type = T_EQ; break;
		case '!':
			if (*p == '=') { p++; t->type = T_NE; }
			else continue;         /* unmatched text is skipped */
			break;
		case '<':
			if (*p == '=') { p++; t->type = T_LE; } else t->type = T_LT;
			break;
		case '>':
			if (*p == '=') { p++; t->type = T_GE; } else t->type = T_GT;
			break;
		case '+': t->type = T_PLUS; break;
		case '-': t->type = T_MINUS; break;
		case '*': t->type = T_MUL; break;
		case '/': t->type = T_DIV; break;
		case ',': t->type = T_COMMA; break;
		case '(': t->type = T_LP; break;
		case ')': t->type = T_RP; break;
		default: continue;         /* unmatched text is skipped */
		}
		n++;
	}
	toks[n].type = T_END;
	*ntok = n + 1;
	return 1;
}

/* -------------------------------------------------------------------- AST */
enum { X_INT = 1, X_FLOAT = 2, X_OP = 3, X_COLUMN = 5 };
enum { XO_PLUS = 1, XO_MINUS = 2, XO_MUL = 3, XO_DIV = 4 };
enum { C_EQ = 1, C_NE = 2, C_LT = 3, C_LE = 4, C_GT = 5, C_GE = 6 };

typedef struct expr {
	int kind;
	int op;            /* X_OP */
	int ival;          /* X_INT, or column id */
	float fval;        /* X_FLOAT */
	int iskey;         /* X_COLUMN */
	char name[64];     /* X_COLUMN */
	virg_t type;
	struct expr *lhs, *rhs;
} expr;

typedef struct cond {
	int cmp;
	expr *lhs, *rhs;
	struct cond *andc, *orc;
	int orfirst;
} cond;

typedef struct rcol {
	expr *e;
	char name[VIRG_MAX_COLUMN_NAME];
	int reg;
} rcol;

#define MAX_NODES 512
#define MAX_RCOLS VIRG_MAX_COLUMNS

typedef struct {
	virginian *v;
	token toks[MAX_TOKENS];
	int ntok, pos;
	expr exprs[MAX_NODES];
	int nexpr;
	cond conds[MAX_NODES];
	int ncond;
	rcol cols[MAX_RCOLS];
	int ncols;
	unsigned table_id;
	cond *where;
	char *err;
	size_t errlen;
	int failed;
} parser;

static void fail(parser *ps, const char *fmt, ...)
{
	if (ps->failed)
		return;
	ps->failed = 1;
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(ps->err, ps->errlen, fmt, ap);
	va_end(ap);
}

static expr *new_expr(parser *ps, int kind)
{
	if (ps->nexpr >= MAX_NODES) {
		fail(ps, "expression too large");
		return NULL;
	}
	expr *e = &ps->exprs[ps->nexpr++];
	memset(e, 0, sizeof(*e));
	e->kind = kind;
	return e;
}

static cond *new_cond(parser *ps)
{
	if (ps->ncond >= MAX_NODES) {
		fail(ps, "too many conditions");
		return NULL;
	}
	cond *c = &ps->conds[ps->ncond++];
	memset(c, 0, sizeof(*c));
	return c;
}

#define TOK (ps->toks[ps->pos].type)

/* expr := primary [op expr]   (right recursive, no precedence) */
static expr *p_expr(parser *ps)
{
	expr *e = NULL;
	token *t = &ps->toks[ps->pos];
	if (t->type == T_LP) {
		ps->pos++;
		e = p_expr(ps);
		if (!e || TOK != T_RP)
			return NULL;
		ps->pos++;
	} else if (t->type == T_INT) {
		e = new_expr(ps, X_INT);
		if (!e) return NULL;
		e->ival = t->ival;
		ps->pos++;
	} else if (t->type == T_FLOAT) {
		e = new_expr(ps, X_FLOAT);
		if (!e) return NULL;
		e->fval = t->fval;
		ps->pos++;
	} else if (t->type == T_NAME) {
		e = new_expr(ps, X_COLUMN);
		if (!e) return NULL;
		strcpy(e->name, t->name);
		ps->pos++;
	} else {
		return NULL;
	}
	int op = 0;
	switch (TOK) {
	case T_PLUS: op = XO_PLUS; break;
	case T_MINUS: op = XO_MINUS; break;
	case T_MUL: op = XO_MUL; break;
	case T_DIV: op = XO_DIV; break;
	default: return e;
	}
	ps->pos++;
	expr *r = p_expr(ps);
	if (!r)
		return NULL;
	expr *o = new_expr(ps, X_OP);
	if (!o) return NULL;
	o->op = op;
	o->lhs = e;
	o->rhs = r;
	return o;
}

static int p_cmp(parser *ps)
{
	int c = 0;
	switch (TOK) {
	case T_EQ: c = C_EQ; break;
	case T_NE: c = C_NE; break;
	case T_LT: c = C_LT; break;
	case T_LE: c = C_LE; break;
	case T_GT: c = C_GT; break;
	case T_GE: c = C_GE; break;
	default: return 0;
	}
	ps->pos++;
	return c;
}

static cond *p_condition(parser *ps)
{
	int save = ps->pos, sexpr = ps->nexpr;
	if (TOK == T_LP) {            /* '(' expr cmp expr ')' */
		ps->pos++;
		expr *l = p_expr(ps);
		int c = l ? p_cmp(ps) : 0;
		expr *r = c ? p_expr(ps) : NULL;
		if (r && TOK == T_RP) {
			ps->pos++;
			cond *k = new_cond(ps);
			if (!k) return NULL;
			k->cmp = c;
			k->lhs = l;
			k->rhs = r;
			return k;
		}
		ps->pos = save;
		ps->nexpr = sexpr;
	}
	expr *l = p_expr(ps);
	int c = l ? p_cmp(ps) : 0;
	expr *r = c ? p_expr(ps) : NULL;
	if (!r)
		return NULL;
	cond *k = new_cond(ps);
	if (!k) return NULL;
	k->cmp = c;
	k->lhs = l;
	k->rhs = r;
	return k;
}

static cond *p_conditions(parser *ps)
{
	int save = ps->pos, sexpr = ps->nexpr, scond = ps->ncond;
	if (TOK == T_LP) {            /* '(' condition (AND|OR) conditions ')' */
		ps->pos++;
		cond *c = p_condition(ps);
		if (c && (TOK == T_AND || TOK == T_OR)) {
			int isor = TOK == T_OR;
			ps->pos++;
			cond *rest = p_conditions(ps);
			if (rest && TOK == T_RP) {
				ps->pos++;
				if (isor) {
					c->orc = rest;
					c->orfirst = 1;
				} else {
					c->andc = rest;
				}
				return c;
			}
		}
		ps->pos = save;
		ps->nexpr = sexpr;
		ps->ncond = scond;
	}
	cond *c = p_condition(ps);
	if (!c)
		return NULL;
	if (TOK == T_AND) {
		ps->pos++;
		c->andc = p_conditions(ps);
		if (!c->andc)
			return NULL;
	} else if (TOK == T_OR) {
		ps->pos++;
		c->orc = p_conditions(ps);
		if (!c->orc)
			return NULL;
	}
	return c;
}

/* name of an unlabelled result column (reference: src/sql/node.c:60-118),
 * truncated to VIRG_MAX_COLUMN_NAME-1 characters */
static void name_append(char *buf, size_t *n, const char *s)
{
	for (; *s && *n < VIRG_MAX_COLUMN_NAME - 1; s++)
		buf[(*n)++] = *s;
	buf[*n] = '\0';
}

static void expr_name(const expr *e, char *buf, size_t *n)
{
	char tmp[64];
	switch (e->kind) {
	case X_COLUMN: name_append(buf, n, e->name); break;
	case X_INT: snprintf(tmp, sizeof tmp, "%i", e->ival); name_append(buf, n, tmp); break;
	case X_FLOAT: snprintf(tmp, sizeof tmp, "%f", e->fval); name_append(buf, n, tmp); break;
	case X_OP: {
		static const char ops[] = " +-*/";
		name_append(buf, n, "(");
		expr_name(e->lhs, buf, n);
		tmp[0] = ops[e->op];
		tmp[1] = '\0';
		name_append(buf, n, tmp);
		expr_name(e->rhs, buf, n);
		name_append(buf, n, ")");
		break;
	}
	}
}

static int p_select(parser *ps)
{
	if (TOK != T_SELECT)
		return 0;
	ps->pos++;
	while (1) {
		if (ps->ncols >= MAX_RCOLS) {
			fail(ps, "too many result columns");
			return 0;
		}
		expr *e = p_expr(ps);
		if (!e)
			return 0;
		rcol *rc = &ps->cols[ps->ncols++];
		rc->e = e;
		rc->name[0] = '\0';
		if (TOK == T_AS && ps->toks[ps->pos + 1].type == T_NAME) {
			ps->pos++;
		}
		if (TOK == T_NAME) {
			snprintf(rc->name, sizeof rc->name, "%s", ps->toks[ps->pos].name);
			ps->pos++;
		} else {
			size_t n = 0;
			expr_name(e, rc->name, &n);
		}
		if (TOK != T_COMMA)
			break;
		ps->pos++;
	}
	if (TOK != T_FROM)
		return 0;
	ps->pos++;
	if (TOK != T_NAME)
		return 0;
	if (virg_table_getid(ps->v, ps->toks[ps->pos].name, &ps->table_id) == VIRG_FAIL) {
		fail(ps, "unknown table '%s'", ps->toks[ps->pos].name);
		return 0;
	}
	ps->pos++;
	ps->where = NULL;
	if (TOK == T_WHERE) {
		ps->pos++;
		ps->where = p_conditions(ps);
		if (!ps->where)
			return 0;
	}
	return TOK == T_END;
}

/* ---------------------------------------------------- pass 0: datatypes */
static int type_expr(parser *ps, expr *e)
{
	switch (e->kind) {
	case X_COLUMN: {
		e->iskey = strcmp(e->name, "id") == 0;
		virg_t t;
		if (e->iskey) {
			e->ival = 0;
			virg_table_getkeytype(ps->v, ps->table_id, &t);
		} else {
			unsigned cid;
			if (virg_table_getcolumn(ps->v, ps->table_id, e->name, &cid) == VIRG_FAIL) {
				fail(ps, "unknown column '%s'", e->name);
				return 0;
			}
			e->ival = (int)cid;
			virg_table_getcolumntype(ps->v, ps->table_id, cid, &t);
		}
		e->type = t;
		return 1;
	}
	case X_OP:
		if (!type_expr(ps, e->lhs) || !type_expr(ps, e->rhs))
			return 0;
		e->type = virg_generalizetype(e->lhs->type, e->rhs->type);
		if (e->type == VIRG_NULL) {
			fail(ps, "operands of incompatible types");
			return 0;
		}
		return 1;
	case X_INT: e->type = VIRG_INT; return 1;
	case X_FLOAT: e->type = VIRG_FLOAT; return 1;
	}
	return 0;
}

static int type_cond(parser *ps, cond *c)
{
	if (!type_expr(ps, c->lhs) || !type_expr(ps, c->rhs))
		return 0;
	if (c->andc && !type_cond(ps, c->andc))
		return 0;
	if (c->orc && !type_cond(ps, c->orc))
		return 0;
	return 1;
}

/* ------------------------------------------------ pass 1: constant fold */
static int as_int(const expr *e)
{
	return e->kind == X_INT ? e->ival : (int)e->fval;
}

static int fold_expr(parser *ps, expr *e)
{
	if (e->kind != X_OP)
		return 1;
	if (!fold_expr(ps, e->lhs) || !fold_expr(ps, e->rhs))
		return 0;
	int lc = e->lhs->kind == X_INT || e->lhs->kind == X_FLOAT;
	int rc = e->rhs->kind == X_INT || e->rhs->kind == X_FLOAT;
	if (!lc || !rc)
		return 1;
	virg_t tt = virg_generalizetype(e->lhs->kind == X_INT ? VIRG_INT : VIRG_FLOAT,
		e->rhs->kind == X_INT ? VIRG_INT : VIRG_FLOAT);
	if (tt == VIRG_INT) {
		int a = as_int(e->lhs), b = as_int(e->rhs), r = 0;
		switch (e->op) {
		case XO_PLUS: r = (int)((unsigned)a + (unsigned)b); break;
		case XO_MINUS: r = (int)((unsigned)a - (unsigned)b); break;
		case XO_MUL: r = (int)((unsigned)a * (unsigned)b); break;
		case XO_DIV:
			if (b == 0 || (a == (-2147483647 - 1) && b == -1)) {
				fail(ps, "integer division by zero or overflow in a constant expression");
				return 0;
			}
			r = a / b;
			break;
		}
		e->kind = X_INT;
		e->ival = r;
	} else {
		/* both operands go through an int conversion first, as the
		 * reference does (generate.c:336-337) */
		float a = (float)as_int(e->lhs), b = (float)as_int(e->rhs), r = 0;
		switch (e->op) {
		case XO_PLUS: r = a + b; break;
		case XO_MINUS: r = a - b; break;
		case XO_MUL: r = a * b; break;
		case XO_DIV: r = a / b; break;
		}
		e->kind = X_FLOAT;
		e->fval = r;
	}
	e->type = tt;
	return 1;
}

static int fold_cond(parser *ps, cond *c)
{
	if (!fold_expr(ps, c->lhs) || !fold_expr(ps, c->rhs))
		return 0;
	if (c->andc && !fold_cond(ps, c->andc))
		return 0;
	if (c->orc && !fold_cond(ps, c->orc))
		return 0;
	return 1;
}

/* ------------------------------------------------ pass 2..5: structure */
typedef struct aop {
	int index;
	virg_op op;
	struct aop *target;   /* jump target / Parallel's converge */
	int is_nop;
} aop;

#define MAX_AOPS 128

typedef struct {
	parser *ps;
	aop ops[MAX_AOPS];
	aop *list[MAX_AOPS];   /* program order */
	int nlist, nops;
	expr *reg_expr[VIRG_REGS];
	int reg_index[VIRG_REGS];
	int nregs;
} gen;

static aop *g_new(gen *g, int op, int p1, int p2, int p3)
{
	if (g->nops >= MAX_AOPS) {
		fail(g->ps, "op-program too long");
		return NULL;
	}
	aop *a = &g->ops[g->nops++];
	memset(a, 0, sizeof(*a));
	a->op.op = op;
	a->op.p1 = p1;
	a->op.p2 = p2;
	a->op.p3 = p3;
	a->op.p4.li = 0;
	a->is_nop = op == OP_Nop;
	return a;
}

static int g_append(gen *g, aop *a)
{
	if (!a)
		return 0;
	if (g->nlist >= MAX_AOPS) {
		fail(g->ps, "op-program too long");
		return 0;
	}
	g->list[g->nlist++] = a;
	return 1;
}

static int expr_equal(const expr *a, const expr *b)
{
	if (a->kind != b->kind)
		return 0;
	switch (a->kind) {
	case X_INT: return a->ival == b->ival;
	case X_FLOAT: return a->fval == b->fval;
	case X_COLUMN: return a->iskey == b->iskey && a->ival == b->ival;
	case X_OP: return a->op == b->op && expr_equal(a->lhs, b->lhs) && expr_equal(a->rhs, b->rhs);
	}
	return 0;
}

static int g_getreg(gen *g)
{
	if (g->nregs >= VIRG_REGS) {
		fail(g->ps, "query needs more than %d registers", VIRG_REGS);
		return -1;
	}
	g->reg_expr[g->nregs] = NULL;
	return g->nregs++;
}

/* reference: src/sql/generate.c:394 */
static int g_expr(gen *g, expr *e)
{
	for (int i = 0; i < g->nregs; i++)
		if (g->reg_expr[i] && expr_equal(e, g->reg_expr[i]))
			return i;
	int reg = -1;
	aop *a = NULL;
	switch (e->kind) {
	case X_INT:
		reg = g_getreg(g);
		if (reg < 0) return -1;
		a = g_new(g, OP_Integer, reg, e->ival, 0);
		break;
	case X_FLOAT:
		reg = g_getreg(g);
		if (reg < 0) return -1;
		a = g_new(g, OP_Float, reg, 0, 0);
		if (a) a->op.p4.f = e->fval;
		break;
	case X_OP: {
		int r1 = g_expr(g, e->lhs);
		int r2 = r1 < 0 ? -1 : g_expr(g, e->rhs);
		if (r2 < 0) return -1;
		reg = g_getreg(g);
		if (reg < 0) return -1;
		static const int mops[] = { 0, OP_Add, OP_Sub, OP_Mul, OP_Div };
		a = g_new(g, mops[e->op], reg, r1, r2);
		break;
	}
	case X_COLUMN:
		reg = g_getreg(g);
		if (reg < 0) return -1;
		a = e->iskey ? g_new(g, OP_Rowid, reg, 0, 0) : g_new(g, OP_Column, reg, e->ival, 0);
		break;
	}
	if (!g_append(g, a))
		return -1;
	g->reg_expr[reg] = e;
	return reg;
}

/* reference: src/sql/generate.c:477 */
static aop *g_cond(gen *g, cond *x, aop *onsuccess, aop *onfailure, aop *newop)
{
	static const int direct[] = { 0, OP_Eq, OP_Neq, OP_Lt, OP_Le, OP_Gt, OP_Ge };
	static const int negated[] = { 0, OP_Neq, OP_Eq, OP_Ge, OP_Gt, OP_Le, OP_Lt };
	int r1 = g_expr(g, x->lhs);
	int r2 = r1 < 0 ? -1 : g_expr(g, x->rhs);
	if (r2 < 0)
		return NULL;
	if (x->andc == NULL || (x->orc != NULL && x->orfirst)) {
		aop *andop = NULL;
		if (x->andc != NULL) {
			andop = g_new(g, OP_Nop, 0, 0, 0);
			onsuccess = andop;
		}
		if (newop == NULL)
			newop = g_new(g, OP_Nop, 0, 0, 0);
		if (!newop) return NULL;
		newop->is_nop = 0;
		newop->op.op = direct[x->cmp];
		newop->op.p1 = r1;
		newop->op.p2 = r2;
		newop->op.p4.li = 0;
		newop->op.p4.i = 1;
		newop->target = onsuccess;
		if (!g_append(g, newop)) return NULL;
		if (x->orc && !g_cond(g, x->orc, onsuccess, onfailure, NULL))
			return NULL;
		if (x->andc && !g_cond(g, x->andc, onfailure, onfailure, andop))
			return NULL;
	} else {
		aop *orop = NULL;
		if (x->orc != NULL) {
			orop = g_new(g, OP_Nop, 0, 0, 0);
			onfailure = orop;
		}
		if (newop == NULL)
			newop = g_new(g, OP_Nop, 0, 0, 0);
		if (!newop) return NULL;
		newop->is_nop = 0;
		newop->op.op = negated[x->cmp];
		newop->op.p1 = r1;
		newop->op.p2 = r2;
		newop->op.p4.li = 0;
		newop->op.p4.i = 0;
		newop->target = onfailure;
		if (!g_append(g, newop)) return NULL;
		if (x->andc && !g_cond(g, x->andc, onsuccess, onfailure, NULL))
			return NULL;
		if (x->orc && !g_cond(g, x->orc, onsuccess, onfailure, orop))
			return NULL;
	}
	return newop;
}

static int generate(parser *ps, virg_vm *vm)
{
	gen *g = (gen *)calloc(1, sizeof(gen));
	if (!g) {
		fail(ps, "out of memory");
		return 0;
	}
	g->ps = ps;
	int ok = 0;

	g_append(g, g_new(g, OP_Table, (int)ps->table_id, 0, 0));
	for (int i = 0; i < ps->ncols; i++) {
		aop *a = g_new(g, OP_ResultColumn, (int)ps->cols[i].e->type, 0, 0);
		if (!a) goto out;
		a->op.p4.s = NULL;
		a->op.p2 = i;  /* name slot, replaced when emitting */
		g_append(g, a);
	}
	aop *result = g_new(g, OP_Result, 0, 0, 0);
	aop *converge = g_new(g, OP_Converge, 0, 0, 0);
	aop *par = g_new(g, OP_Parallel, 0, 0, 0);
	if (!result || !converge || !par) goto out;
	par->target = converge;
	g_append(g, par);

	if (ps->where) {
		aop *stub = g_new(g, OP_Nop, 0, 0, 0);
		if (!stub) goto out;
		if (!g_cond(g, ps->where, stub, result, NULL)) goto out;
		g_append(g, g_new(g, OP_Invalid, 0, 0, 0));
		g_append(g, stub);
	}
	for (int i = 0; i < ps->ncols; i++) {
		int r = g_expr(g, ps->cols[i].e);
		if (r < 0) goto out;
		ps->cols[i].reg = r;
	}
	/* move result registers to the back, in result-column order */
	for (int i = 0; i < g->nregs; i++)
		g->reg_index[i] = i;
	for (int c = 0; c < ps->ncols; c++) {
		int r = ps->cols[c].reg;
		if (g->reg_index[r] == g->nregs - 1)
			continue;
		int old = g->reg_index[r];
		g->reg_index[r] = g->nregs - 1;
		for (int i = 0; i < g->nregs; i++)
			if (i != r && g->reg_index[i] > old)
				g->reg_index[i]--;
	}
	result->op.p1 = ps->cols[0].reg;
	result->op.p2 = ps->ncols;
	g_append(g, result);
	g_append(g, converge);
	g_append(g, g_new(g, OP_Finish, 0, 0, 0));
	if (ps->failed) goto out;

	/* pass 3: indices (a Nop takes the index of the op after it) */
	for (int i = 0, idx = 0; i < g->nlist; i++)
		g->list[i]->index = g->list[i]->is_nop ? idx : idx++;
	/* pass 4: registers and jumps */
	for (int i = 0; i < g->nlist; i++) {
		aop *a = g->list[i];
		switch (a->op.op) {
		case OP_Parallel: a->op.p3 = a->target->index + 1; break;
		case OP_Integer: case OP_Column: case OP_Rowid: case OP_Result: case OP_Float:
			a->op.p1 = g->reg_index[a->op.p1];
			break;
		case OP_Add: case OP_Sub: case OP_Mul: case OP_Div:
			a->op.p1 = g->reg_index[a->op.p1];
			a->op.p2 = g->reg_index[a->op.p2];
			a->op.p3 = g->reg_index[a->op.p3];
			break;
		case OP_Eq: case OP_Neq: case OP_Le: case OP_Lt: case OP_Ge: case OP_Gt:
			a->op.p1 = g->reg_index[a->op.p1];
			a->op.p2 = g->reg_index[a->op.p2];
			a->op.p3 = a->target->index;
			break;
		default: break;
		}
	}
	/* pass 5: emit */
	vm->num_ops = 0;
	for (int i = 0; i < g->nlist; i++) {
		aop *a = g->list[i];
		if (a->is_nop)
			continue;
		virg_var p4 = a->op.p4;
		int p2 = a->op.p2;
		if (a->op.op == OP_ResultColumn) {
			p4.s = (char *)malloc(VIRG_MAX_COLUMN_NAME);
			if (!p4.s) {
				fail(ps, "out of memory");
				goto out;
			}
			strcpy(p4.s, ps->cols[a->op.p2].name);
			p2 = 0;
		}
		if (virg_vm_addop(vm, a->op.op, a->op.p1, p2, a->op.p3, p4) == VIRG_FAIL) {
			if (a->op.op == OP_ResultColumn)
				free(p4.s);
			fail(ps, "op-program longer than %d ops", VIRG_OPS);
			goto out;
		}
	}
	ok = 1;
out:
	free(g);
	return ok;
}

int virg_sql_compile(virginian *v, const char *querystr, virg_vm *vm, char *err, size_t errlen)
{
	parser *ps = (parser *)calloc(1, sizeof(parser));
	if (!ps) {
		snprintf(err, errlen, "out of memory");
		return VIRG_FAIL;
	}
	ps->v = v;
	ps->err = err;
	ps->errlen = errlen;
	err[0] = '\0';
	int ok = lex(querystr, ps->toks, &ps->ntok, err, errlen);
	if (ok && !p_select(ps)) {
		if (!ps->failed)
			fail(ps, "Parse Error :: syntax error near token %d", ps->pos);
		ok = 0;
	}
	if (ok) {
		for (int i = 0; i < ps->ncols && ok; i++)
			ok = type_expr(ps, ps->cols[i].e);
		if (ok && ps->where)
			ok = type_cond(ps, ps->where);
		for (int i = 0; i < ps->ncols && ok; i++)
			ok = fold_expr(ps, ps->cols[i].e);
		if (ok && ps->where)
			ok = fold_cond(ps, ps->where);
		if (ok)
			ok = generate(ps, vm);
	}
	free(ps);
	return ok ? VIRG_SUCCESS : VIRG_FAIL;
}

/* reference: src/sql/sql.y:228 */
int virg_sql(virginian *v, const char *querystr, virg_vm *vm)
{
	char err[256];
	if (virg_sql_compile(v, querystr, vm, err, sizeof err) == VIRG_FAIL) {
		fprintf(stderr, "::: Virginian SQL error: %s\n", err);
		return VIRG_FAIL;
	}
	return VIRG_SUCCESS;
}
