/*
 * Tablet layout and growth.
 *
 * A tablet is one contiguous block:
 *   [virg_tablet_meta (704 B)][keys: key_stride * possible_rows]
 *   [key pointers: 8 * possible_rows][fixed columns, column-major,
 *   column j holding fixed_stride[j] * possible_rows bytes][variable]
 * (reference: src/tablet/create.c:52-60, src/tablet/addcolumn.c:36-37,
 * src/tablet/addrows.c:39-73, src/tablet/addtail.c:37-50).  Standard
 * tablets are capped at VIRG_TABLET_SIZE like the reference; the bulk
 * builder (virg_table_appendtablet) makes tablets of any row capacity with
 * the same layout, which is how 1M-row tablets are built.
 */
#include "internal.h"
#include <stdlib.h>
#include <string.h>

size_t virg_sizeof(virg_t type)
{
	return virg_sizes[type];
}

/* reference: src/util.c:14-26 */
virg_t virg_generalizetype(virg_t t1, virg_t t2)
{
	if (t1 == t2)
		return t1;
	if (t1 <= VIRG_DOUBLE && t2 <= VIRG_DOUBLE)
		return t1 > t2 ? t1 : t2;
	return VIRG_NULL;
}

/* reference: src/tablet/create.c:19 */
int virg_tablet_create(virginian *v, int *id, virg_t key_type, unsigned table_id)
{
	virg_tablet_meta *m;
	int tid = (int)v->db.tablet_id_counter++;
	VIRG_CHECK(virg_db_alloc(v, &m, tid) == VIRG_FAIL, "Could not allocate tablet");
	m->rows = 0;
	m->key_type = key_type;
	m->key_stride = virg_sizeof(key_type);
	m->key_pointer_stride = sizeof(size_t);
	m->row_stride = m->key_stride + m->key_pointer_stride;
	m->last_tablet = 1;
	m->id = (unsigned)tid;
	id[0] = tid;
	if (table_id > VIRG_MAX_TABLES) {
		m->in_table = 0;
		m->table_id = 0;
	} else {
		m->in_table = 1;
		m->table_id = table_id;
	}
	m->possible_rows = VIRG_TABLET_INITIAL_KEYS;
	m->key_block = sizeof(virg_tablet_meta);
	m->key_pointers_block = m->key_block + m->key_stride * VIRG_TABLET_INITIAL_KEYS;
	m->fixed_block = m->key_pointers_block + m->key_pointer_stride * VIRG_TABLET_INITIAL_KEYS;
	m->variable_block = m->fixed_block + VIRG_TABLET_INITIAL_FIXED;
	m->size = m->variable_block + VIRG_TABLET_INITIAL_VARIABLE;
	m->info = NULL;
	m->fixed_columns = 0;
	virg_tablet_unlock(v, (unsigned)tid);
	return VIRG_SUCCESS;
}

/* reference: src/tablet/growfixed.c:19 */
int virg_tablet_growfixed(virg_tablet_meta *tab, size_t size)
{
	VIRG_CHECK(tab->size + size > VIRG_TABLET_SIZE, "Too big to grow, the tablet must be split");
	size_t var_size = tab->size - tab->variable_block;
	if (var_size > 0)
		memmove((char *)tab + tab->variable_block + size,
			(char *)tab + tab->variable_block, var_size);
	tab->variable_block += size;
	tab->size += size;
	return VIRG_SUCCESS;
}

/* reference: src/tablet/addcolumn.c:19 */
int virg_tablet_addcolumn(virg_tablet_meta *tab, const char *name, virg_t type)
{
	unsigned col = tab->fixed_columns;
	VIRG_CHECK(col == VIRG_MAX_COLUMNS, "Too many columns");
	size_t n = strlen(name);
	if (n >= VIRG_MAX_COLUMN_NAME)
		n = VIRG_MAX_COLUMN_NAME - 1;
	memcpy(tab->fixed_name[col], name, n);
	tab->fixed_name[col][n] = '\0';
	tab->fixed_type[col] = type;
	tab->fixed_stride[col] = virg_sizeof(type);
	tab->row_stride += virg_sizeof(type);
	tab->fixed_offset[col] = col == 0 ? 0 :
		tab->fixed_offset[col - 1] + tab->fixed_stride[col - 1] * tab->possible_rows;
	tab->fixed_columns++;
	return virg_tablet_growfixed(tab, tab->fixed_stride[col] * tab->possible_rows);
}

/* Re-lay a tablet for `new_possible` rows in place: fixed columns move to
 * their new offsets (back to front, since they only move up), then the key
 * pointer block.  Shared by addrows and the result-tablet sizing. */
static void relayout(virg_tablet_meta *tab, unsigned new_possible)
{
	size_t offs[VIRG_MAX_COLUMNS];
	size_t new_kp = tab->key_block + tab->key_stride * new_possible;
	size_t new_fixed = new_kp + tab->key_pointer_stride * new_possible;
	for (unsigned i = 0; i < tab->fixed_columns; i++)
		offs[i] = i == 0 ? 0 : offs[i - 1] + tab->fixed_stride[i - 1] * new_possible;
	for (unsigned i = tab->fixed_columns; i-- > 0;)
		memmove((char *)tab + new_fixed + offs[i],
			(char *)tab + tab->fixed_block + tab->fixed_offset[i],
			tab->rows * tab->fixed_stride[i]);
	memmove((char *)tab + new_kp, (char *)tab + tab->key_pointers_block,
		tab->rows * tab->key_pointer_stride);
	for (unsigned i = 0; i < tab->fixed_columns; i++)
		tab->fixed_offset[i] = offs[i];
	tab->fixed_block = new_fixed;
	tab->key_pointers_block = new_kp;
	tab->possible_rows = new_possible;
}

/* reference: src/tablet/addrows.c:22.  Rounds to multiples of 16 rows;
 * rows that do not fit go to new tail tablets. */
int virg_tablet_addrows(virginian *v, virg_tablet_meta *tab, unsigned rows)
{
	size_t rs = tab->row_stride;
	unsigned max_new = (unsigned)((VIRG_TABLET_SIZE - tab->size) / rs) & 0xFFFFFFF0u;
	rows = (rows - 1 + 16) & 0xFFFFFFF0u;
	unsigned new_rows = VIRG_MIN(max_new, rows);
	if (new_rows != 0) {
		VIRG_CHECK(virg_tablet_growfixed(tab, rs * new_rows) == VIRG_FAIL, "grow failed");
		relayout(tab, tab->possible_rows + new_rows);
	}
	if (max_new <= rows) {
		virg_tablet_lock(v, tab->id);
		tab->size = VIRG_TABLET_SIZE;
		unsigned left = rows - new_rows;
		unsigned max_tab = (unsigned)((VIRG_TABLET_SIZE - sizeof(virg_tablet_meta) -
			VIRG_TABLET_INITIAL_FIXED) / rs);
		virg_tablet_meta *node = tab, *tail = tab;
		while (left > 0) {
			unsigned r = VIRG_MIN(left, max_tab);
			VIRG_CHECK(virg_tablet_addtail(v, node, &tail, r) == VIRG_FAIL, "addtail failed");
			left -= r;
			node = tail;
		}
		if (tab->in_table) {
			v->db.last_tablet[tab->table_id] = tail->id;
			v->db.table_tablets[tab->table_id]++;
		}
		virg_tablet_unlock(v, node->id);
	}
	return VIRG_SUCCESS;
}

/* reference: src/tablet/addmaxrows.c:15 */
int virg_tablet_addmaxrows(virginian *v, virg_tablet_meta *tab)
{
	size_t avail = VIRG_TABLET_SIZE - sizeof(virg_tablet_meta) - VIRG_TABLET_MAXED_VARIABLE;
	unsigned pr = (unsigned)(avail / tab->row_stride);
	return virg_tablet_addrows(v, tab, pr - tab->rows);
}

/* reference: src/tablet/addtail.c:18.  The new tail comes back locked;
 * head is unlocked. */
int virg_tablet_addtail(virginian *v, virg_tablet_meta *head,
	virg_tablet_meta **tail, unsigned possible_rows)
{
	virg_tablet_meta *m;
	int tid = (int)v->db.tablet_id_counter++;
	VIRG_CHECK(virg_db_alloc(v, &m, tid) == VIRG_FAIL, "Could not allocate tail tablet");
	memcpy(m, head, sizeof(virg_tablet_meta));
	head->last_tablet = 0;
	head->next = (unsigned)tid;
	virg_tablet_unlock(v, head->id);
	tail[0] = m;
	m->rows = 0;
	m->id = (unsigned)tid;
	m->key_pointers_block = m->key_block + m->key_stride * possible_rows;
	m->fixed_block = m->key_pointers_block + m->key_pointer_stride * possible_rows;
	for (unsigned i = 1; i < m->fixed_columns; i++)
		m->fixed_offset[i] = m->fixed_offset[i - 1] + m->fixed_stride[i - 1] * possible_rows;
	m->variable_block = m->key_block + m->row_stride * possible_rows;
	m->size = m->variable_block + VIRG_TABLET_INITIAL_VARIABLE;
	m->info = NULL;
	m->possible_rows = possible_rows;
	return VIRG_SUCCESS;
}

/* reference: src/tablet/check.c:17 -- layout invariants */
int virg_tablet_check(virg_tablet_meta *t)
{
	VIRG_CHECK(t->rows > t->possible_rows, "more rows than capacity");
	VIRG_CHECK(t->key_block != sizeof(virg_tablet_meta), "key block misplaced");
	VIRG_CHECK(t->key_pointers_block != t->key_block + t->key_stride * t->possible_rows,
		"key pointer block misplaced");
	VIRG_CHECK(t->fixed_block != t->key_pointers_block + t->key_pointer_stride * t->possible_rows,
		"fixed block misplaced");
	size_t stride = t->key_stride + t->key_pointer_stride;
	for (unsigned i = 0; i < t->fixed_columns; i++) {
		size_t want = i == 0 ? 0 : t->fixed_offset[i - 1] + t->fixed_stride[i - 1] * t->possible_rows;
		VIRG_CHECK(t->fixed_offset[i] != want, "column offset wrong");
		VIRG_CHECK(t->fixed_stride[i] != virg_sizeof(t->fixed_type[i]), "column stride wrong");
		stride += t->fixed_stride[i];
	}
	VIRG_CHECK(stride != t->row_stride, "row stride wrong");
	VIRG_CHECK(t->variable_block < t->fixed_block + (t->row_stride - t->key_stride -
		t->key_pointer_stride) * t->possible_rows, "variable block overlaps columns");
	return VIRG_SUCCESS;
}
