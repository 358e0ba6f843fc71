/*
 * ORACLE -- test infrastructure only.  Never linked into the product.
 *
 * A small driver around the reference's own CPU VM, compiled from the
 * reference sources where they lie (oracle/build_ref.sh) into
 * oracle/_ref/libvirgref.so.  It exists to check the C restatement in
 * virg_oracle.c against the code it restates, on the same inputs.
 *
 * What it replaces, and why:
 *   - src/init.c: its CUDA context and pinned-slot calls need libcudart,
 *     which is not in the image; ref_open() does the rest of virg_init
 *     (defaults, slot mutex, 64 slots of VIRG_TABLET_SIZE with malloc, as
 *     init.c's VIRG_NOPINNED branch does);
 *   - src/vm/execute.c: it also calls virg_vm_gpu (vm_gpu.cu needs nvcc);
 *     ref_run() walks the same serial ops (Table, ResultColumn, Parallel,
 *     Finish; execute.c:28-102) and calls the reference's virg_vm_cpu for
 *     the data-parallel section;
 *   - src/close.c: freeing what ref_open allocated.
 * Everything else -- tablets, table inserts, slot manager, result tablets,
 * the single- and multi-core VM (virginia.c, vm_cpu.c) -- is the
 * reference's code.
 */
#include "virginian.h"

#define API __attribute__((visibility("default")))

void virginia_single(virginian *v, virg_vm *vm, virg_tablet_meta *tab,
	virg_tablet_meta **res, unsigned start_row, unsigned num_rows);
int virg_vm_cpu(virginian *v, virg_vm *vm, virg_tablet_meta **tab, virg_tablet_meta **res,
	unsigned num_tablets);
int virg_db_close(virginian *v);

static virginian *V;

static int ref_init(int use_multi, int threads)
{
	int i;
	V = calloc(1, sizeof(virginian));
	if (!V)
		return VIRG_FAIL;
	V->tablet_slot_counter = 0;
	V->tablet_slots_taken = 0;
	V->threads_per_block = VIRG_THREADSPERBLOCK;
	V->multi_threads = threads > 0 ? (unsigned)threads : VIRG_MULTITHREADS;
	V->use_multi = use_multi;
	V->use_gpu = 0;
	V->use_stream = 0;
	V->use_mmap = 0;
	V->dbfd = -1;
	if (pthread_mutex_init(&V->slot_lock, NULL))
		return VIRG_FAIL;
	for (i = 0; i < VIRG_MEM_TABLETS; i++) {
		V->tablet_slot_status[i] = 0;
		V->tablet_slots[i] = malloc(VIRG_TABLET_SIZE);
		if (!V->tablet_slots[i])
			return VIRG_FAIL;
	}
	return VIRG_SUCCESS;
}

/* A new database backed by a new file (the reference's virg_db_create) */
API int ref_open(const char *dbfile, int use_multi, int threads)
{
	if (ref_init(use_multi, threads) != VIRG_SUCCESS)
		return VIRG_FAIL;
	unlink(dbfile);
	return virg_db_create(V, dbfile);
}

/* An existing database file (the reference's virg_db_open); tablets are
 * read from it on first use by virg_db_load */
API int ref_open_file(const char *dbfile, int use_multi, int threads)
{
	if (ref_init(use_multi, threads) != VIRG_SUCCESS)
		return VIRG_FAIL;
	return virg_db_open(V, dbfile);
}

API void ref_close(void)
{
	int i;
	if (!V)
		return;
	if (V->dbfd >= 0)
		close(V->dbfd);
	for (i = 0; i < VIRG_MEM_TABLETS; i++)
		free(V->tablet_slots[i]);
	free(V->db.tablet_info);
	pthread_mutex_destroy(&V->slot_lock);
	free(V);
	V = NULL;
}

/* Close as src/close.c:20 does: virg_db_close writes every tablet in a slot
 * and the meta information to the file; then free what ref_init took. */
API int ref_db_close(void)
{
	int ok;
	if (!V)
		return VIRG_FAIL;
	ok = virg_db_close(V);
	V->db.tablet_info = NULL;   /* freed by virg_db_close */
	ref_close();
	return ok;
}

API int ref_table(const char *name, int key_type, int ncols, const char *const *names,
	const int *types)
{
	unsigned id;
	int j;
	if (virg_table_create(V, name, (virg_t)key_type) == VIRG_FAIL)
		return -1;
	if (virg_table_getid(V, name, &id) == VIRG_FAIL)
		return -1;
	for (j = 0; j < ncols; j++)
		if (virg_table_addcolumn(V, id, names[j], (virg_t)types[j]) == VIRG_FAIL)
			return -1;
	return (int)id;
}

/* Insert rows one at a time through virg_table_insert (the reference's
 * only ingest path): key[r] and, per column, cols[j][r]. */
API int ref_insert(unsigned table, const void *key, int key_width, int ncols,
	const void *const *cols, const int *widths, unsigned long long rows)
{
	char kbuf[8], row[VIRG_MAX_COLUMNS * 8];
	unsigned long long r;
	for (r = 0; r < rows; r++) {
		int j, off = 0;
		memcpy(kbuf, (const char *)key + r * key_width, key_width);
		for (j = 0; j < ncols; j++) {
			memcpy(row + off, (const char *)cols[j] + r * widths[j], widths[j]);
			off += widths[j];
		}
		if (virg_table_insert(V, table, kbuf, row, NULL) == VIRG_FAIL)
			return VIRG_FAIL;
	}
	return VIRG_SUCCESS;
}

/* Run a program through the serial ops of execute.c and the reference's
 * virg_vm_cpu; returns the vm with its result-tablet chain, or NULL. */
static virg_vm *ref_exec(const virg_op *ops, unsigned num_ops)
{
	virg_vm *vm = virg_vm_init();
	virg_tablet_meta *tab = NULL, *res;
	unsigned i;
	for (i = 0; i < num_ops; i++) {
		virg_var p4 = ops[i].p4;
		/* virg_vm_cleanup frees ResultColumn names (vm/cleanup.c:17-22) */
		if (ops[i].op == OP_ResultColumn) {
			const char *nm = ops[i].p4.s ? ops[i].p4.s : "";
			p4.s = malloc(strlen(nm) + 1);
			if (!p4.s)
				return NULL;
			memcpy(p4.s, nm, strlen(nm) + 1);
		}
		if (virg_vm_addop(vm, ops[i].op, ops[i].p1, ops[i].p2, ops[i].p3, p4) == VIRG_FAIL)
			return NULL;
	}
	/* execute.c:38-44 */
	vm->num_tables = 0;
	vm->pc = 0;
	vm->head_result = NULL;
	virg_vm_allocresult(V, vm, &res, NULL);
	for (;;) {
		const virg_op *o = &vm->stmt[vm->pc];
		if (o->op == OP_Table) {                 /* execute.c:63-70 */
			vm->table[vm->num_tables++] = o->p1;
			virg_db_load(V, V->db.first_tablet[o->p1], &tab);
			vm->pc++;
		} else if (o->op == OP_ResultColumn) {   /* :72-75 */
			virg_tablet_addcolumn(res, o->p4.s, (virg_t)o->p1);
			vm->pc++;
		} else if (o->op == OP_Parallel) {      /* :77-90 */
			int p3 = o->p3;
			virg_tablet_addmaxrows(V, res);
			vm->pc++;
			virg_vm_cpu(V, vm, &tab, &res, 0);
			vm->pc = p3;
		} else if (o->op == OP_Finish) {        /* :92-96 */
			virg_tablet_unlock(V, tab->id);
			virg_tablet_unlock(V, res->id);
			break;
		} else {
			return NULL;
		}
	}
	return vm;
}

/* Run a program; copy the logical result (the result-tablet chain in
 * order) into out[j] up to cap rows.  Returns the row count or -1. */
API long long ref_run(const virg_op *ops, unsigned num_ops, void **out, unsigned nout,
	unsigned long long cap)
{
	virg_vm *vm = ref_exec(ops, num_ops);
	long long total = 0;
	if (!vm)
		return -1;
	/* the logical result: every result tablet of the chain, in order */
	virg_result_node *node = vm->head_result;
	while (node) {
		virg_tablet_meta *r;
		virg_db_load(V, node->id, &r);
		unsigned j;
		for (j = 0; j < nout && j < r->fixed_columns; j++) {
			unsigned long long n = r->rows;
			if ((unsigned long long)total + n > cap)
				n = cap - total;
			memcpy((char *)out[j] + (size_t)total * r->fixed_stride[j],
				(char *)r + r->fixed_block + r->fixed_offset[j], n * r->fixed_stride[j]);
		}
		total += r->rows;
		virg_tablet_unlock(V, r->id);
		node = node->next;
	}
	virg_vm_freeresults(V, vm);
	virg_vm_cleanup(V, vm);
	return total;
}

/* Run a program as ref_run does, then read its result with the reference's
 * reader (src/reader/init.c, row.c): rows + k * row_bytes receives
 * r.buffer after the k-th call of virg_reader_row, for every call up to
 * and including the first that returns VIRG_FAIL (at most cap calls);
 * *nsuccess = the calls that returned VIRG_SUCCESS.  Returns the calls
 * made, or -1. */
API long long ref_run_reader(const virg_op *ops, unsigned num_ops, void *rows, unsigned row_bytes,
	unsigned long long cap, unsigned long long *nsuccess)
{
	virg_vm *vm = ref_exec(ops, num_ops);
	virg_reader r;
	unsigned long long calls = 0;
	if (!vm || row_bytes > VIRG_ROW_BUFFER)
		return -1;
	*nsuccess = 0;
	if (virg_reader_init(V, &r, vm) != VIRG_SUCCESS)
		return -1;
	while (calls < cap) {
		const int rc = virg_reader_row(V, &r);
		memcpy((char *)rows + calls * row_bytes, r.buffer, row_bytes);
		calls++;
		if (rc == VIRG_FAIL)
			break;
		(*nsuccess)++;
	}
	virg_reader_free(V, &r);
	virg_vm_freeresults(V, vm);
	virg_vm_cleanup(V, vm);
	return (long long)calls;
}
