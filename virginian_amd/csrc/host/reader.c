/*
 * Result cursor over the chain of result tablets (reference: src/reader/).
 * The logical result sequence is the concatenation of the result tablets
 * from vm->head_result in chain order.
 */
#include "internal.h"
#include <stdlib.h>
#include <string.h>

/* reference: src/reader/init.c:18 */
int virg_reader_init(virginian *v, virg_reader *r, virg_vm *vm)
{
	r->vm = vm;
	r->res = NULL;
	r->row = 0;
	VIRG_CHECK(vm->head_result == NULL, "No results");
	return virg_db_load(v, vm->head_result->id, &r->res);
}

/* reference: src/reader/free.c:17 */
int virg_reader_free(virginian *v, virg_reader *r)
{
	if (r->res != NULL)
		virg_tablet_unlock(v, r->res->id);
	r->res = NULL;
	return VIRG_SUCCESS;
}

/* reference: src/reader/row.c:20.  Copies the current row into r->buffer
 * and advances; like the reference it returns VIRG_FAIL after copying the
 * final row of the last tablet. */
int virg_reader_row(virginian *v, virg_reader *r)
{
	if (r->res == NULL)
		return VIRG_FAIL;
	virg_tablet_meta *res = r->res;
	char *dest = r->buffer;
	char *fixed = (char *)res + res->fixed_block;
	for (unsigned i = 0; i < res->fixed_columns; i++) {
		size_t s = res->fixed_stride[i];
		memcpy(dest, fixed + res->fixed_offset[i] + s * r->row, s);
		dest += s;
	}
	r->row++;
	if (r->row >= res->rows) {
		if (res->last_tablet) {
			virg_tablet_unlock(v, res->id);
			r->res = NULL;
			return VIRG_FAIL;
		}
		virg_db_loadnext(v, &r->res);
		r->row = 0;
	}
	return VIRG_SUCCESS;
}

/* reference: src/reader/nexttablet.c:17 */
int virg_reader_nexttablet(virginian *v, virg_reader *r)
{
	VIRG_CHECK(r->res->last_tablet, "Reached the last tablet");
	r->row = 0;
	return virg_db_loadnext(v, &r->res);
}

/* reference: src/reader/getrows.c:19 */
int virg_reader_getrows(virginian *v, virg_reader *r, unsigned *rows_)
{
	if (r->res == NULL)
		return VIRG_FAIL;
	virg_tablet_meta *res = r->res;
	unsigned rows = 0;
	virg_tablet_lock(v, res->id);
	while (1) {
		rows += res->rows;
		if (res->last_tablet)
			break;
		virg_db_loadnext(v, &res);
	}
	virg_tablet_unlock(v, res->id);
	rows_[0] = rows - r->row;
	return VIRG_SUCCESS;
}

/* Copy the whole logical result of vm into caller arrays, one per result
 * column (cols[j] receives rows * fixed_stride[j] bytes).  Returns the row
 * count in *rows; with cols == NULL only counts.  No reference
 * counterpart: the reference reads results a row at a time. */
int virg_result_columns(virginian *v, virg_vm *vm, void **cols, unsigned long long *rows)
{
	virg_result_node *node = vm->head_result;
	unsigned long long n = 0;
	for (; node != NULL; node = node->next) {
		virg_tablet_meta *res;
		VIRG_CHECK(virg_db_load(v, node->id, &res) == VIRG_FAIL, "missing result tablet");
		if (cols)
			for (unsigned j = 0; j < res->fixed_columns; j++) {
				size_t s = res->fixed_stride[j];
				memcpy((char *)cols[j] + n * s,
					(char *)res + res->fixed_block + res->fixed_offset[j],
					(size_t)res->rows * s);
			}
		n += res->rows;
		virg_tablet_unlock(v, res->id);
	}
	rows[0] = n;
	return VIRG_SUCCESS;
}

/* reference: src/query.c:21 */
int virg_query(virginian *v, virg_reader **reader, const char *query)
{
	reader[0] = NULL;
	virg_vm *vm = virg_vm_init();
	VIRG_CHECK(vm == NULL, "Out of memory");
	if (virg_sql(v, query, vm) == VIRG_FAIL || virg_vm_execute(v, vm) == VIRG_FAIL) {
		virg_vm_cleanup(v, vm);
		return VIRG_FAIL;
	}
	virg_reader *r = (virg_reader *)malloc(sizeof(virg_reader));
	if (r == NULL || virg_reader_init(v, r, vm) == VIRG_FAIL) {
		free(r);
		virg_vm_cleanup(v, vm);
		return VIRG_FAIL;
	}
	reader[0] = r;
	return VIRG_SUCCESS;
}

/* reference: src/release.c:17 */
void virg_release(virginian *v, virg_reader *reader)
{
	virg_reader_free(v, reader);
	virg_vm_cleanup(v, reader->vm);
	free(reader);
}
