/*
 * Top-level VM: context lifecycle, result tablets and the serial
 * interpreter of Table / ResultColumn / Parallel / Finish.
 *
 * The data-parallel section (Parallel .. Converge) always runs on the GPU
 * through virg_vm_gpu: this library has no CPU data-parallel VM, so a
 * query with use_gpu == 0 fails instead of silently running on the host.
 */
#include "internal.h"
#include <stdlib.h>
#include <string.h>

/* reference: src/vm/init.c:13 */
virg_vm *virg_vm_init(void)
{
	virg_vm *vm = (virg_vm *)calloc(1, sizeof(virg_vm));
	if (!vm)
		return NULL;
	vm->pc = 0;
	vm->head_result = NULL;
	vm->tail_result = NULL;
	vm->num_ops = 0;
	vm->num_tables = 0;
	return vm;
}

/* reference: src/vm/addop.c:19 */
int virg_vm_addop(virg_vm *vm, int op, int p1, int p2, int p3, virg_var p4)
{
	VIRG_CHECK(vm->num_ops >= VIRG_OPS, "Too many ops");
	virg_op *o = &vm->stmt[vm->num_ops++];
	o->op = op;
	o->p1 = p1;
	o->p2 = p2;
	o->p3 = p3;
	o->p4 = p4;
	return VIRG_SUCCESS;
}

/* reference: src/vm/allocresult.c:22.  A new result tablet either starts
 * empty (key INT, no columns, no capacity) or copies the template's meta,
 * and is appended to vm's result list and the template's chain. */
int virg_vm_allocresult(virginian *v, virg_vm *vm, virg_tablet_meta **meta,
	virg_tablet_meta *template_)
{
	virg_tablet_meta *tab;
	unsigned id = v->db.tablet_id_counter++;
	size_t bytes = template_ ? sizeof(virg_tablet_meta) + template_->size : VIRG_TABLET_SIZE;
	if (bytes < VIRG_TABLET_SIZE)
		bytes = VIRG_TABLET_SIZE;
	VIRG_CHECK(virg_db_alloc_sized(v, &tab, (int)id, bytes) == VIRG_FAIL,
		"Could not allocate result tablet");
	if (template_ == NULL) {
		tab->key_type = VIRG_INT;
		tab->key_stride = virg_sizeof(VIRG_INT);
		tab->last_tablet = 1;
		tab->key_pointer_stride = sizeof(size_t);
		tab->key_block = sizeof(virg_tablet_meta);
		tab->possible_rows = 0;
		tab->key_pointers_block = tab->key_block;
		tab->fixed_block = tab->key_pointers_block;
		tab->variable_block = tab->fixed_block;
		tab->size = tab->variable_block;
		tab->row_stride = tab->key_stride + tab->key_pointer_stride;
		tab->fixed_columns = 0;
		tab->info = NULL;
		tab->in_table = 0;
	} else {
		memcpy(tab, template_, sizeof(virg_tablet_meta));
		template_->last_tablet = 0;
		tab->id = id;
		template_->next = id;
	}
	tab->rows = 0;
	tab->next = 0;

	virg_result_node *node = (virg_result_node *)malloc(sizeof(virg_result_node));
	VIRG_CHECK(node == NULL, "Out of memory");
	node->id = id;
	node->next = NULL;
	if (vm->head_result == NULL)
		vm->head_result = node;
	else
		vm->tail_result->next = node;
	vm->tail_result = node;
	meta[0] = tab;
	return VIRG_SUCCESS;
}

/* A result tablet like virg_vm_allocresult's from `template_`, laid out for
 * `rows` rows instead of the template's capacity (the reference layout,
 * tablet/addcolumn.c:36-37, applied to `rows`): the device path gives the
 * rest of a large result one tablet instead of many 8 MB ones. */
int virg_vm_allocresult_rows(virginian *v, virg_vm *vm, virg_tablet_meta **meta,
	virg_tablet_meta *template_, unsigned rows)
{
	virg_tablet_meta t = *template_;
	t.possible_rows = rows;
	t.key_block = sizeof(virg_tablet_meta);
	t.key_pointers_block = t.key_block + t.key_stride * rows;
	t.fixed_block = t.key_pointers_block + t.key_pointer_stride * rows;
	for (unsigned c = 0; c < t.fixed_columns; c++)
		t.fixed_offset[c] = c == 0 ? 0 : t.fixed_offset[c - 1] + t.fixed_stride[c - 1] * rows;
	t.variable_block = t.key_block + t.row_stride * rows;
	t.size = t.variable_block;
	const size_t bytes = sizeof(virg_tablet_meta) + t.size;
	unsigned id = v->db.tablet_id_counter++;
	virg_tablet_meta *tab;
	VIRG_CHECK(virg_db_alloc_sized(v, &tab, (int)id, bytes < VIRG_TABLET_SIZE ? VIRG_TABLET_SIZE : bytes) == VIRG_FAIL,
		"Could not allocate result tablet");
	*tab = t;
	template_->last_tablet = 0;
	template_->next = id;
	tab->id = id;
	tab->rows = 0;
	tab->next = 0;
	tab->last_tablet = 1;
	tab->info = NULL;
	virg_result_node *node = (virg_result_node *)malloc(sizeof(virg_result_node));
	VIRG_CHECK(node == NULL, "Out of memory");
	node->id = id;
	node->next = NULL;
	if (vm->head_result == NULL)
		vm->head_result = node;
	else
		vm->tail_result->next = node;
	vm->tail_result = node;
	meta[0] = tab;
	return VIRG_SUCCESS;
}

/* reference: src/vm/freeresults.c:17 */
int virg_vm_freeresults(virginian *v, virg_vm *vm)
{
	virg_result_node *node = vm->head_result;
	while (node != NULL) {
		virg_result_node *next = node->next;
		virg_tablet_remove(v, node->id);
		free(node);
		node = next;
	}
	vm->head_result = NULL;
	vm->tail_result = NULL;
	return VIRG_SUCCESS;
}

/* reference: src/vm/cleanup.c:13 */
void virg_vm_cleanup(virginian *v, virg_vm *vm)
{
	if (vm->head_result != NULL)
		virg_vm_freeresults(v, vm);
	for (unsigned i = 0; i < vm->num_ops; i++)
		if (vm->stmt[i].op == OP_ResultColumn)
			free(vm->stmt[i].p4.s);
	free(vm);
}

/* reference: src/vm/execute.c:28 */
int virg_vm_execute(virginian *v, virg_vm *vm)
{
	virg_tablet_meta *tab = NULL, *res = NULL;
	int rc = VIRG_SUCCESS;
	vm->num_tables = 0;
	vm->pc = 0;
	vm->head_result = NULL;
	vm->tail_result = NULL;
	VIRG_CHECK(virg_vm_allocresult(v, vm, &res, NULL) == VIRG_FAIL, "result allocation failed");

	while (vm->pc < vm->num_ops) {
		virg_op *op = &vm->stmt[vm->pc];
		switch (op->op) {
		case OP_Table:
			VIRG_CHECK(vm->num_tables >= VIRG_VM_TABLES, "Too many tables in query");
			VIRG_CHECK(op->p1 < 0 || op->p1 >= VIRG_MAX_TABLES ||
				!v->db.table_status[op->p1], "Query names an unknown table");
			vm->table[vm->num_tables++] = (unsigned)op->p1;
			VIRG_CHECK(virg_db_load(v, v->db.first_tablet[op->p1], &tab) == VIRG_FAIL,
				"could not load first tablet");
			vm->pc++;
			break;
		case OP_ResultColumn:
			VIRG_CHECK(virg_tablet_addcolumn(res, op->p4.s, (virg_t)op->p1) == VIRG_FAIL,
				"could not add result column");
			vm->pc++;
			break;
		case OP_Parallel: {
			unsigned after = (unsigned)op->p3;
			VIRG_CHECK(tab == NULL, "Parallel before Table");
			virg_tablet_addmaxrows(v, res);
			vm->pc++;
			if (!v->use_gpu) {
				VIRG_ERROR("this library executes the data-parallel VM on the GPU only (set use_gpu)");
				rc = VIRG_FAIL;
			} else if (virg_vm_gpu(v, vm, &tab, &res, 0) == VIRG_FAIL) {
				rc = VIRG_FAIL;
			}
			vm->pc = after;
			break;
		}
		case OP_Finish:
			if (tab)
				virg_tablet_unlock(v, tab->id);
			virg_tablet_unlock(v, res->id);
			return rc;
		default:
			VIRG_ERROR("Invalid OP");
			return VIRG_FAIL;
		}
	}
	VIRG_ERROR("op-program has no Finish");
	return VIRG_FAIL;
}

static const char *const opnames[] = {
	"Table", "ResultColumn", "Parallel", "Finish", "Column", "Rowid", "Result",
	"Converge", "Invalid", "Cast", "Integer", "Float", "Le", "Lt", "Ge", "Gt",
	"Eq", "Neq", "Add", "Sub", "Mul", "Div", "And", "Or", "Not", "Nop"
};

/* reference: src/opcodelist.awk (generated name table) */
const char *virg_opstring(int op)
{
	if (op < 0 || op > OP_Nop)
		return "";
	return opnames[op];
}

struct timeval virg_starttime;
struct timeval virg_endtime;

/* reference: src/util.c:51-75 */
void virg_timer_start(void)
{
	gettimeofday(&virg_starttime, NULL);
}

double virg_timer_stop(void)
{
	gettimeofday(&virg_endtime, NULL);
	return (double)(virg_endtime.tv_sec - virg_starttime.tv_sec) +
		(double)(virg_endtime.tv_usec - virg_starttime.tv_usec) * 1e-6;
}
