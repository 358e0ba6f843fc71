/*
 * Decoder: virg_vm data-parallel section -> vg_prog.
 *
 * Semantics preserved (reference CPU VM, src/vm/virginia.c):
 *  - Column/Rowid load the column for every row of a block, whatever the
 *    row's pc (:442-474), so a register defined once by Column is the
 *    column itself and needs no micro-op.
 *  - Integer/Float/arithmetic write rows whose pc is the current op
 *    (:386-412, :138-194).  A register written once is written for every
 *    row that can later read it in a well-formed program, so the first
 *    write is done for all rows; later writes of the same register keep
 *    the masked semantics (VG_F_MASKED).
 *  - Compares: a row at this op whose compare is true jumps to p3 and, if
 *    valid, takes validity p4 (:48-130); jumps are forward only.  A jump
 *    past Result lands on Result (Result emits every valid row, :476-578).
 *  - Register types are per register, not per row (:449, :144): resolved
 *    here, with the reference's requirement that compare and arithmetic
 *    operands have one type (:52, :143).
 * The section must end Result, Converge (what the SQL generator emits,
 * reference: src/sql/generate.c:676-682).
 */
#include "virginian.h"
#include "vg_prog.h"
#include <stdarg.h>
#include <string.h>

enum { RC_NONE = 0, RC_CONST, RC_COLUMN, RC_SLOT };

/* The last micro-op after which a slot may be given to another value.
 * Built with VG_DECODE_PRIVATE_SLOTS, no slot is ever given twice: every
 * value keeps a slot of its own, the layout the tests check the reusing
 * one against. */
#ifdef VG_DECODE_PRIVATE_SLOTS
#define HELD_UNTIL(k) (VG_MAX_UOPS + 1)
#else
#define HELD_UNTIL(k) (k)
#endif

typedef struct {
	int cls;
	int ndefs;
	int type;           /* type at the point of use, updated in program order */
	int max_width;
	int col_src;        /* RC_COLUMN */
	uint64_t imm;       /* RC_CONST */
	int first_def_uop;  /* RC_SLOT live range */
	int last_use_uop;
	int slot, slot_hi;
} regstate;

static int derr(char *err, size_t n, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(err, n, fmt, ap);
	va_end(ap);
	return VIRG_FAIL;
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

static int cmp_kind(int op)
{
	switch (op) {
	case OP_Lt: return VG_K_LT;
	case OP_Le: return VG_K_LE;
	case OP_Gt: return VG_K_GT;
	case OP_Ge: return VG_K_GE;
	case OP_Eq: return VG_K_EQ;
	case OP_Neq: return VG_K_NE;
	case OP_And: return VG_K_AND;
	case OP_Or: return VG_K_OR;
	}
	return -1;
}

static int math_kind(int op)
{
	switch (op) {
	case OP_Add: return VG_K_ADD;
	case OP_Sub: return VG_K_SUB;
	case OP_Mul: return VG_K_MUL;
	case OP_Div: return VG_K_DIV;
	}
	return -1;
}

/* operand reference recorded during the first pass */
typedef struct {
	int reg;     /* -1 if unused */
} opnd;

typedef struct {
	int kind, type, flags;
	int dst_reg;            /* -1 if none */
	opnd a, b;
	int cast_to;
	int target_op;          /* branches: op index */
	int src_op;
	int from_col;           /* MOV from a column: column source, else -1 */
	uint64_t a_imm_override; int a_is_imm;  /* MOV of an immediate */
} puop;

int vg_decode(const virg_op *stmt, unsigned num_ops, unsigned pc,
	const virg_tablet_meta *tab, const virg_tablet_meta *res,
	vg_prog *P, char *err, size_t errlen)
{
	regstate R[VIRG_REGS];
	puop U[VG_MAX_UOPS];
	int nu = 0;
	memset(R, 0, sizeof R);
	memset(P, 0, sizeof *P);
	for (int r = 0; r < VIRG_REGS; r++) {
		R[r].type = -1;
		R[r].first_def_uop = -1;
		R[r].last_use_uop = -1;
		R[r].slot = R[r].slot_hi = -1;
	}

	/* locate Result / Converge */
	unsigned res_op = pc;
	while (res_op < num_ops && stmt[res_op].op != OP_Result && stmt[res_op].op != OP_Converge)
		res_op++;
	if (res_op >= num_ops)
		return derr(err, errlen, "data-parallel section has no Converge");
	int has_result = stmt[res_op].op == OP_Result;
	if (has_result && (res_op + 1 >= num_ops || stmt[res_op + 1].op != OP_Converge))
		return derr(err, errlen, "Result must be followed by Converge");

	/* pass 1: types and definition counts */
	int type[VIRG_REGS];
	for (int r = 0; r < VIRG_REGS; r++)
		type[r] = -1;
#define CHKREG(x) do { if ((x) < 0 || (x) >= VIRG_REGS) \
	return derr(err, errlen, "op %u: register %d out of range", i, (x)); } while (0)
#define CHKDEF(x) do { CHKREG(x); if (type[(x)] < 0) \
	return derr(err, errlen, "op %u: register %d read before it is written", i, (x)); } while (0)
	for (unsigned i = pc; i < res_op; i++) {
		const virg_op *o = &stmt[i];
		switch (o->op) {
		case OP_Column:
			CHKREG(o->p1);
			if (o->p2 < 0 || (unsigned)o->p2 >= tab->fixed_columns)
				return derr(err, errlen, "op %u: column %d does not exist", i, o->p2);
			type[o->p1] = tab->fixed_type[o->p2];
			R[o->p1].ndefs++;
			break;
		case OP_Rowid:
			CHKREG(o->p1);
			type[o->p1] = tab->key_type;
			R[o->p1].ndefs++;
			break;
		case OP_Integer:
			CHKREG(o->p1);
			type[o->p1] = VIRG_INT;
			R[o->p1].ndefs++;
			break;
		case OP_Float:
			CHKREG(o->p1);
			type[o->p1] = VIRG_FLOAT;
			R[o->p1].ndefs++;
			break;
		case OP_Add: case OP_Sub: case OP_Mul: case OP_Div:
			CHKREG(o->p1);
			CHKDEF(o->p2);
			CHKDEF(o->p3);
			if (type[o->p2] != type[o->p3])
				return derr(err, errlen, "op %u: arithmetic on registers of different types", i);
			type[o->p1] = type[o->p2];
			R[o->p1].ndefs++;
			break;
		case OP_Le: case OP_Lt: case OP_Ge: case OP_Gt: case OP_Eq: case OP_Neq:
		case OP_And: case OP_Or:
			CHKDEF(o->p1);
			CHKDEF(o->p2);
			if (type[o->p1] != type[o->p2])
				return derr(err, errlen, "op %u: compare of registers of different types", i);
			if (o->p3 <= (int)i || (unsigned)o->p3 >= num_ops)
				return derr(err, errlen, "op %u: jump target %d is not a later op", i, o->p3);
			break;
		case OP_Not:
			CHKDEF(o->p1);
			if (o->p3 <= (int)i || (unsigned)o->p3 >= num_ops)
				return derr(err, errlen, "op %u: jump target %d is not a later op", i, o->p3);
			break;
		case OP_Cast:
			CHKDEF(o->p2);
			if (tsize(o->p1) == 0)
				return derr(err, errlen, "op %u: cast to unsupported type %d", i, o->p1);
			type[o->p2] = o->p1;
			R[o->p2].ndefs++;
			break;
		case OP_Invalid:
		case OP_Nop:
			break;
		default:
			return derr(err, errlen, "op %u: %s is not allowed in the data-parallel section",
				i, virg_opstring(o->op));
		}
	}
	for (int r = 0; r < VIRG_REGS; r++)
		if (type[r] >= 0 && tsize(type[r]) == 0)
			return derr(err, errlen, "register %d has unsupported type %d", r, type[r]);

	/* register classes */
	for (unsigned i = pc; i < res_op; i++) {
		const virg_op *o = &stmt[i];
		int r = -1;
		if (o->op == OP_Column || o->op == OP_Rowid || o->op == OP_Integer || o->op == OP_Float)
			r = o->p1;
		if (r < 0 || R[r].ndefs != 1)
			continue;
		if (o->op == OP_Integer) {
			R[r].cls = RC_CONST;
			R[r].imm = (uint64_t)(uint32_t)o->p2;
		} else if (o->op == OP_Float) {
			uint32_t bits;
			memcpy(&bits, &o->p4.f, 4);
			R[r].cls = RC_CONST;
			R[r].imm = bits;
		} else {
			R[r].cls = RC_COLUMN;
			R[r].col_src = o->op == OP_Rowid ? VG_KEY_SRC : o->p2;
		}
	}
	for (int r = 0; r < VIRG_REGS; r++)
		if (R[r].ndefs > 0 && R[r].cls == RC_NONE)
			R[r].cls = RC_SLOT;

	/* pass 2: micro-ops in program order */
	int seen_def[VIRG_REGS];
	memset(seen_def, 0, sizeof seen_def);
	for (int r = 0; r < VIRG_REGS; r++)
		type[r] = -1;
	for (unsigned i = pc; i < res_op; i++) {
		const virg_op *o = &stmt[i];
		puop u;
		memset(&u, 0, sizeof u);
		u.dst_reg = -1;
		u.a.reg = u.b.reg = -1;
		u.from_col = -1;
		u.src_op = (int)i;
		int emit = 1;
		switch (o->op) {
		case OP_Column:
		case OP_Rowid:
			type[o->p1] = o->op == OP_Rowid ? (int)tab->key_type : (int)tab->fixed_type[o->p2];
			if (R[o->p1].cls != RC_SLOT) {
				emit = 0;
				break;
			}
			u.kind = VG_K_MOV;
			u.type = type[o->p1];
			u.dst_reg = o->p1;
			u.from_col = o->op == OP_Rowid ? VG_KEY_SRC : o->p2;
			break;
		case OP_Integer:
		case OP_Float:
			type[o->p1] = o->op == OP_Integer ? VIRG_INT : VIRG_FLOAT;
			if (R[o->p1].cls != RC_SLOT) {
				emit = 0;
				break;
			}
			u.kind = VG_K_MOV;
			u.type = type[o->p1];
			u.dst_reg = o->p1;
			u.a_is_imm = 1;
			if (o->op == OP_Integer) {
				u.a_imm_override = (uint64_t)(uint32_t)o->p2;
			} else {
				uint32_t bits;
				memcpy(&bits, &o->p4.f, 4);
				u.a_imm_override = bits;
			}
			if (seen_def[o->p1])
				u.flags |= VG_F_MASKED;
			break;
		case OP_Add: case OP_Sub: case OP_Mul: case OP_Div:
			u.kind = math_kind(o->op);
			u.type = type[o->p2];
			u.dst_reg = o->p1;
			u.a.reg = o->p2;
			u.b.reg = o->p3;
			if (seen_def[o->p1])
				u.flags |= VG_F_MASKED;
			type[o->p1] = type[o->p2];
			break;
		case OP_Le: case OP_Lt: case OP_Ge: case OP_Gt: case OP_Eq: case OP_Neq:
		case OP_And: case OP_Or:
			u.kind = cmp_kind(o->op);
			u.type = type[o->p1];
			u.a.reg = o->p1;
			u.b.reg = o->p2;
			u.target_op = o->p3 > (int)res_op ? (int)res_op : o->p3;
			if (o->p4.i != 0)
				u.flags |= VG_F_KEEP;
			break;
		case OP_Not:
			u.kind = VG_K_NOT;
			u.type = type[o->p1];
			u.a.reg = o->p1;
			u.target_op = o->p3 > (int)res_op ? (int)res_op : o->p3;
			if (o->p4.i != 0)
				u.flags |= VG_F_KEEP;
			break;
		case OP_Invalid:
			u.kind = VG_K_INVALID;
			break;
		case OP_Cast:
			u.kind = VG_K_CAST;
			u.type = type[o->p2];
			u.cast_to = o->p1;
			u.dst_reg = o->p2;
			u.a.reg = o->p2;
			u.flags |= VG_F_MASKED;
			type[o->p2] = o->p1;
			break;
		default:
			emit = 0;
			break;
		}
		if (o->op == OP_Column || o->op == OP_Rowid || o->op == OP_Integer ||
			o->op == OP_Float || math_kind(o->op) >= 0 || o->op == OP_Cast)
			seen_def[o->op == OP_Cast ? o->p2 : o->p1] = 1;
		if (!emit)
			continue;
		if (nu >= VG_MAX_UOPS)
			return derr(err, errlen, "program needs more than %d micro-ops", VG_MAX_UOPS);
		U[nu++] = u;
	}

	/* validity-changing prefix */
	int nfilter = 0;
	for (int k = 0; k < nu; k++)
		if (U[k].kind <= VG_K_INVALID)
			nfilter = k + 1;

	/* column usage: first micro-op reading each source, and emit use */
	int first_read[VG_MAX_COLS], emitted[VG_MAX_COLS];
	for (int c = 0; c < VG_MAX_COLS; c++) {
		first_read[c] = -1;
		emitted[c] = 0;
	}
	for (int k = 0; k < nu; k++) {
		int srcs[3] = { -1, -1, -1 };
		if (U[k].from_col >= 0)
			srcs[0] = U[k].from_col;
		if (U[k].a.reg >= 0 && R[U[k].a.reg].cls == RC_COLUMN)
			srcs[1] = R[U[k].a.reg].col_src;
		if (U[k].b.reg >= 0 && R[U[k].b.reg].cls == RC_COLUMN)
			srcs[2] = R[U[k].b.reg].col_src;
		for (int j = 0; j < 3; j++)
			if (srcs[j] >= 0 && first_read[srcs[j]] < 0)
				first_read[srcs[j]] = k;
	}
	unsigned nres = has_result ? (unsigned)stmt[res_op].p2 : 0;
	int res_p1 = has_result ? stmt[res_op].p1 : 0;
	if (nres > VG_MAX_OUT)
		return derr(err, errlen, "Result emits %u registers, at most %d supported", nres, VG_MAX_OUT);
	for (unsigned j = 0; j < nres; j++) {
		int r = res_p1 + (int)j;
		if (r < 0 || r >= VIRG_REGS || R[r].cls == RC_NONE)
			return derr(err, errlen, "Result register %d is never written", r);
		if (R[r].cls == RC_COLUMN)
			emitted[R[r].col_src] = 1;
	}

	/* column table: filter loads, project loads, then emit-only */
	int col_index[VG_MAX_COLS];
	for (int c = 0; c < VG_MAX_COLS; c++)
		col_index[c] = -1;
	int nc = 0, next_slot = 0;
	for (int pass = 0; pass < 3; pass++)
		for (int c = 0; c < VG_MAX_COLS; c++) {
			int phase;
			if (first_read[c] >= 0)
				phase = first_read[c] < nfilter ? VG_PHASE_FILTER : VG_PHASE_PROJECT;
			else if (emitted[c])
				phase = VG_PHASE_EMIT;
			else
				continue;
			if (phase != pass)
				continue;
			vg_col *col = &P->col[nc];
			int ty = c == VG_KEY_SRC ? (int)tab->key_type : (int)tab->fixed_type[c];
			col->src = (uint8_t)c;
			col->type = (uint8_t)ty;
			col->width = (uint8_t)tsize(ty);
			col->phase = (uint8_t)phase;
			if (phase == VG_PHASE_FILTER) {
				/* filter columns take the lowest slots so the count pass,
				 * which only touches filter slots, needs the least LDS */
				int need = col->width == 8 ? 2 : 1;
				if (next_slot + need > VG_NSLOT)
					return derr(err, errlen, "computation reads too many columns");
				col->slot = (uint8_t)next_slot;
				col->slot_hi = (uint8_t)(next_slot + need - 1);
				next_slot += need;
				P->nload = (uint32_t)(nc + 1);
			} else if (phase == VG_PHASE_PROJECT) {
				col->slot = col->slot_hi = 0xFF;   /* placed after the temporaries */
				P->nload = (uint32_t)(nc + 1);
				P->has_project_loads = 1;
			}
			col_index[c] = nc++;
		}
	P->ncol = (uint32_t)nc;

	/* live ranges of slot registers */
	for (int k = 0; k < nu; k++) {
		int regs[2] = { U[k].a.reg, U[k].b.reg };
		for (int j = 0; j < 2; j++) {
			int r = regs[j];
			if (r >= 0 && R[r].cls == RC_SLOT)
				R[r].last_use_uop = k;
		}
		if (U[k].dst_reg >= 0) {
			regstate *s = &R[U[k].dst_reg];
			if (s->first_def_uop < 0)
				s->first_def_uop = k;
			if (s->last_use_uop < k)
				s->last_use_uop = k;
			int w = tsize(U[k].kind == VG_K_CAST ? U[k].cast_to : U[k].type);
			if (w > s->max_width)
				s->max_width = w;
		}
	}
	for (unsigned j = 0; j < nres; j++) {
		int r = res_p1 + (int)j;
		if (R[r].cls == RC_SLOT)
			R[r].last_use_uop = nu;
	}

	/* linear-scan slot allocation in micro-op order */
	int slot_free_at[VG_NSLOT];
	for (int s = 0; s < VG_NSLOT; s++)
		slot_free_at[s] = -1;    /* last micro-op using the slot */
	for (int k = 0; k < nu; k++) {
		int r = U[k].dst_reg;
		if (r < 0 || R[r].first_def_uop != k)
			continue;
		int need = R[r].max_width == 8 ? 2 : 1;
		int got[2] = { -1, -1 }, ng = 0;
		for (int s = next_slot; s < VG_NSLOT && ng < need; s++)
			if (slot_free_at[s] < k)
				got[ng++] = s;
		if (ng < need)
			return derr(err, errlen, "program needs more than %d register slots", VG_NSLOT);
		R[r].slot = got[0];
		R[r].slot_hi = need == 2 ? got[1] : got[0];
		for (int j = 0; j < ng; j++)
			slot_free_at[got[j]] = HELD_UNTIL(R[r].last_use_uop);
	}

	/* slots the count pass touches: filter columns and the values the
	 * filter micro-ops write */
	int nslot_count = next_slot;
	for (int k = 0; k < nfilter; k++) {
		int r = U[k].dst_reg;
		if (r >= 0 && R[r].slot_hi + 1 > nslot_count)
			nslot_count = R[r].slot_hi + 1;
	}
	/* projection columns go above every temporary */
	int top = next_slot;
	for (int r = 0; r < VIRG_REGS; r++)
		if (R[r].cls == RC_SLOT && R[r].slot_hi + 1 > top)
			top = R[r].slot_hi + 1;
	for (int c = 0; c < nc; c++) {
		vg_col *col = &P->col[c];
		if (col->phase != VG_PHASE_PROJECT)
			continue;
		int need = col->width == 8 ? 2 : 1;
		if (top + need > VG_NSLOT)
			return derr(err, errlen, "program needs more than %d register slots", VG_NSLOT);
		col->slot = (uint8_t)top;
		col->slot_hi = (uint8_t)(top + need - 1);
		top += need;
	}
	P->nslot_count = (uint32_t)nslot_count;

	/* final micro-ops */
	int maxslot = top;
	for (int k = 0; k < nu; k++) {
		vg_uop *o = &P->uop[k];
		puop *u = &U[k];
		o->kind = (uint8_t)u->kind;
		o->type = (uint8_t)u->type;
		o->flags = (uint8_t)u->flags;
		o->src_op = (uint16_t)u->src_op;
		if (u->kind <= VG_K_NOT) {
			int t = nu;
			for (int j = 0; j < nu; j++)
				if (U[j].src_op >= u->target_op) {
					t = j;
					break;
				}
			o->target = (uint8_t)t;
		}
		if (u->dst_reg >= 0) {
			o->dst = (uint8_t)R[u->dst_reg].slot;
			o->dst_hi = (uint8_t)R[u->dst_reg].slot_hi;
			if (R[u->dst_reg].slot_hi + 1 > maxslot)
				maxslot = R[u->dst_reg].slot_hi + 1;
		}
		if (u->kind == VG_K_CAST)
			o->b = (uint8_t)u->cast_to;
		if (u->from_col >= 0) {
			const vg_col *c = &P->col[col_index[u->from_col]];
			o->a = c->slot;
			o->a_hi = c->slot_hi;
		} else if (u->a_is_imm) {
			o->flags |= VG_F_AIMM;
			o->ia = u->a_imm_override;
		}
		opnd *ops[2] = { &u->a, &u->b };
		for (int j = 0; j < 2; j++) {
			int r = ops[j]->reg;
			if (r < 0)
				continue;
			uint8_t lo, hi;
			uint64_t imm = 0;
			int isimm = 0;
			if (R[r].cls == RC_CONST) {
				isimm = 1;
				imm = R[r].imm;
				lo = hi = 0;
			} else if (R[r].cls == RC_COLUMN) {
				const vg_col *c = &P->col[col_index[R[r].col_src]];
				lo = c->slot;
				hi = c->slot_hi;
			} else {
				lo = (uint8_t)R[r].slot;
				hi = (uint8_t)R[r].slot_hi;
			}
			if (j == 0) {
				if (isimm) { o->flags |= VG_F_AIMM; o->ia = imm; }
				else { o->a = lo; o->a_hi = hi; }
			} else {
				if (isimm) { o->flags |= VG_F_BIMM; o->ib = imm; }
				else { o->b = lo; o->b_hi = hi; }
			}
		}
	}
	P->nuop = (uint32_t)nu;
	P->nfilter = (uint32_t)nfilter;

	/* Wait slots.  A row that takes a branch keeping its validity skips
	 * the micro-ops up to the target (the reference's row_pc jumps
	 * forward, virginia.c:59-62) and resumes there; rows that lose
	 * validity never reach the result again, so only keep-branches need a
	 * slot.  Branches with the same target share one; a slot is free
	 * again once its target is reached. */
	{
		int slot_target[VG_NWAIT];
		for (int j = 0; j < VG_NWAIT; j++)
			slot_target[j] = -1;
		for (int k = 0; k < nu; k++) {
			vg_uop *o = &P->uop[k];
			for (int j = 0; j < VG_NWAIT; j++)
				if (slot_target[j] == k) {
					o->join |= (uint8_t)(1u << j);
					slot_target[j] = -1;
				}
			if (o->kind > VG_K_NOT || !(o->flags & VG_F_KEEP))
				continue;
			int s = -1;
			for (int j = 0; j < VG_NWAIT && s < 0; j++)
				if (slot_target[j] == (int)o->target)
					s = j;
			for (int j = 0; j < VG_NWAIT && s < 0; j++)
				if (slot_target[j] < 0)
					s = j;
			if (s < 0)
				return derr(err, errlen, "op %d: more than %d forward-jump targets pending at once",
					o->src_op, VG_NWAIT);
			o->wslot = (uint8_t)s;
			slot_target[s] = o->target;
		}
	}

	/* outputs */
	for (unsigned j = 0; j < nres; j++) {
		int r = res_p1 + (int)j;
		vg_out *out = &P->out[j];
		int ty = type[r];
		out->type = (uint8_t)ty;
		out->width = (uint8_t)tsize(ty);
		if (res != NULL) {
			if (j >= res->fixed_columns)
				return derr(err, errlen, "Result emits %u registers but the result has %u columns",
					nres, res->fixed_columns);
			if (res->fixed_stride[j] != out->width)
				return derr(err, errlen, "result column %u holds %zu-byte values, register %d is %d bytes",
					j, res->fixed_stride[j], r, out->width);
		}
		if (R[r].cls == RC_CONST) {
			out->kind = VG_OUT_CONST;
			out->imm = R[r].imm;
		} else if (R[r].cls == RC_COLUMN) {
			const vg_col *c = &P->col[col_index[R[r].col_src]];
			if (c->phase == VG_PHASE_EMIT) {
				out->kind = VG_OUT_COLUMN;
				out->col = (uint8_t)col_index[R[r].col_src];
			} else {
				out->kind = VG_OUT_SLOT;
				out->slot = c->slot;
				out->slot_hi = c->slot_hi;
			}
		} else {
			out->kind = VG_OUT_SLOT;
			out->slot = (uint8_t)R[r].slot;
			out->slot_hi = (uint8_t)R[r].slot_hi;
		}
	}
	P->nout = nres;
	P->nslot = (uint32_t)maxslot;

	/* Write pass: validity is already known from the count pass, so it
	 * only needs the micro-ops that compute result registers (and what
	 * they read).  Masked writes depend on per-row resume points, which
	 * only a full rerun reproduces. */
	int full = 0;
	for (int k = 0; k < nu; k++)
		if (P->uop[k].flags & VG_F_MASKED)
			full = 1;
	uint64_t need = 0;
	uint8_t slot_live[VG_NSLOT];
	memset(slot_live, 0, sizeof slot_live);
	for (unsigned j = 0; j < nres; j++)
		if (P->out[j].kind == VG_OUT_SLOT) {
			slot_live[P->out[j].slot] = 1;
			slot_live[P->out[j].slot_hi] = 1;
		}
	if (full) {
		need = nu >= 64 ? ~0ull : ((1ull << nu) - 1);
	} else {
		/* backward over micro-ops: a value op is needed if it writes a
		 * live slot; its operands then become live */
		for (int k = nu - 1; k >= 0; k--) {
			const vg_uop *o = &P->uop[k];
			if (o->kind <= VG_K_INVALID)
				continue;
			if (!slot_live[o->dst] && !slot_live[o->dst_hi])
				continue;
			need |= 1ull << k;
			slot_live[o->dst] = slot_live[o->dst_hi] = 0;
			if (!(o->flags & VG_F_AIMM)) {
				slot_live[o->a] = 1;
				slot_live[o->a_hi] = 1;
			}
			if (o->kind != VG_K_MOV && o->kind != VG_K_CAST && !(o->flags & VG_F_BIMM)) {
				slot_live[o->b] = 1;
				slot_live[o->b_hi] = 1;
			}
		}
	}
	P->wfull = (uint32_t)full;
	P->wneed = need;
	for (unsigned c = 0; c < P->nload; c++) {
		vg_col *col = &P->col[c];
		int used = full;
		for (int k = 0; k < nu && !used; k++) {
			if (!((need >> k) & 1))
				continue;
			const vg_uop *o = &P->uop[k];
			if (!(o->flags & VG_F_AIMM) && (o->a == col->slot || o->a_hi == col->slot))
				used = 1;
			if (o->kind != VG_K_MOV && o->kind != VG_K_CAST && !(o->flags & VG_F_BIMM) &&
				(o->b == col->slot || o->b_hi == col->slot))
				used = 1;
		}
		for (unsigned j = 0; j < nres && !used; j++)
			if (P->out[j].kind == VG_OUT_SLOT && P->out[j].slot == col->slot)
				used = 1;
		col->wload = (uint8_t)used;
	}
	P->wsafe = 1;
	for (unsigned j = 0; j < nres; j++) {
		const vg_out *o = &P->out[j];
		if (o->kind != VG_OUT_SLOT)
			continue;
		for (unsigned c = 0; c < P->nload; c++)
			if (P->col[c].wload && (o->slot == P->col[c].slot || o->slot == P->col[c].slot_hi ||
					o->slot_hi == P->col[c].slot || o->slot_hi == P->col[c].slot_hi))
				P->wsafe = 0;
	}
	return VIRG_SUCCESS;
}
