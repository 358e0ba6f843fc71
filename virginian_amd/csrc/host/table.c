/*
 * Tables: a named chain of tablets sharing a schema
 * (reference: src/table/).  Row-at-a-time insert follows the reference
 * (src/table/insert.c:38-88); virg_table_appendtablet is the bulk path that
 * builds large tablets directly from column arrays.
 */
#include "internal.h"
#include <stdlib.h>
#include <string.h>

/* reference: src/table/create.c:18 */
int virg_table_create(virginian *v, const char *name, virg_t key_type)
{
	virg_db *db = &v->db;
	unsigned t;
	for (t = 0; t < VIRG_MAX_TABLES; t++)
		VIRG_CHECK(db->table_status[t] && strcmp(db->tables[t], name) == 0, "Table exists");
	for (t = 0; t < VIRG_MAX_TABLES; t++)
		if (db->table_status[t] == 0)
			break;
	VIRG_CHECK(t == VIRG_MAX_TABLES, "Too many tables");
	VIRG_CHECK(strlen(name) >= VIRG_MAX_TABLE_NAME, "Table name too long");
	db->table_status[t] = 1;
	strcpy(db->tables[t], name);
	int tid;
	VIRG_CHECK(virg_tablet_create(v, &tid, key_type, t) == VIRG_FAIL, "tablet create failed");
	db->first_tablet[t] = (unsigned)tid;
	db->last_tablet[t] = (unsigned)tid;
	db->write_cursor[t] = (unsigned)tid;
	db->table_tablets[t]++;
	return VIRG_SUCCESS;
}

/* reference: src/table/addcolumn.c:18 */
int virg_table_addcolumn(virginian *v, unsigned table_id, const char *name, virg_t type)
{
	VIRG_CHECK(strlen(name) >= VIRG_MAX_COLUMN_NAME, "Column name too long");
	virg_tablet_meta *tab;
	VIRG_CHECK(virg_db_load(v, v->db.first_tablet[table_id], &tab) == VIRG_FAIL, "no tablet");
	while (1) {
		VIRG_CHECK(virg_tablet_addcolumn(tab, name, type) == VIRG_FAIL, "add column failed");
		if (tab->last_tablet)
			break;
		virg_db_loadnext(v, &tab);
	}
	virg_tablet_unlock(v, tab->id);
	return VIRG_SUCCESS;
}

/* reference: src/table/insert.c:38 */
int virg_table_insert(virginian *v, unsigned table_id, char *key, char *data, char *blob)
{
	virg_tablet_meta *tab;
	VIRG_CHECK(blob != NULL, "blob columns are not supported");
	VIRG_CHECK(virg_db_load(v, v->db.write_cursor[table_id], &tab) == VIRG_FAIL, "no tablet");
	if (tab->rows == tab->possible_rows) {
		if (tab->size < VIRG_TABLET_SIZE - tab->row_stride)
			virg_tablet_addrows(v, tab, VIRG_TABLET_KEY_INCREMENT);
		else {
			virg_db_loadnext(v, &tab);
			v->db.write_cursor[tab->table_id] = tab->id;
		}
	}
	char *fixed = (char *)tab + tab->fixed_block;
	memcpy((char *)tab + tab->key_block + (size_t)tab->rows * tab->key_stride, key, tab->key_stride);
	const char *src = data;
	for (unsigned i = 0; i < tab->fixed_columns; i++) {
		size_t s = tab->fixed_stride[i];
		memcpy(fixed + tab->fixed_offset[i] + (size_t)tab->rows * s, src, s);
		src += s;
	}
	tab->rows++;
	virg_tablet_unlock(v, tab->id);
	return VIRG_SUCCESS;
}

int virg_table_appendtablet(virginian *v, unsigned table_id, unsigned capacity,
	unsigned rows, const void *key, const void *const *cols)
{
	VIRG_CHECK(table_id >= VIRG_MAX_TABLES || !v->db.table_status[table_id], "no such table");
	VIRG_CHECK(rows > capacity || capacity == 0, "rows exceed tablet capacity");
	virg_tablet_meta *first, *last;
	VIRG_CHECK(virg_db_load(v, v->db.first_tablet[table_id], &first) == VIRG_FAIL, "no tablet");
	VIRG_CHECK(virg_db_load(v, v->db.last_tablet[table_id], &last) == VIRG_FAIL, "no tablet");

	virg_tablet_meta schema = *first;
	size_t bytes = sizeof(virg_tablet_meta) + (size_t)capacity * schema.row_stride;
	virg_tablet_meta *m;
	int reuse = (last->rows == 0 && last->id == first->id && first->last_tablet);
	int tid;
	if (reuse) {
		/* the table's initial empty tablet is replaced in place: free its
		 * slot and allocate the big one under the same id */
		tid = (int)last->id;
		virg_tablet_unlock(v, last->id);
		virg_tablet_unlock(v, first->id);
		virg_tablet_remove(v, (unsigned)tid);
	} else {
		tid = (int)v->db.tablet_id_counter++;
	}
	VIRG_CHECK(virg_db_alloc_sized(v, &m, tid, bytes) == VIRG_FAIL, "Could not allocate tablet");
	*m = schema;
	m->id = (unsigned)tid;
	m->rows = rows;
	m->next = 0;
	m->last_tablet = 1;
	m->in_table = 1;
	m->table_id = table_id;
	m->info = NULL;
	m->possible_rows = capacity;
	m->key_block = sizeof(virg_tablet_meta);
	m->key_pointers_block = m->key_block + m->key_stride * capacity;
	m->fixed_block = m->key_pointers_block + m->key_pointer_stride * capacity;
	for (unsigned i = 0; i < m->fixed_columns; i++)
		m->fixed_offset[i] = i == 0 ? 0 : m->fixed_offset[i - 1] + m->fixed_stride[i - 1] * capacity;
	m->variable_block = m->key_block + m->row_stride * capacity;
	m->size = m->variable_block;
	if (rows) {
		/* with no key or columns given the caller fills them in place */
		if (key)
			memcpy((char *)m + m->key_block, key, m->key_stride * rows);
		memset((char *)m + m->key_pointers_block, 0, m->key_pointer_stride * rows);
		for (unsigned i = 0; cols && i < m->fixed_columns; i++)
			memcpy((char *)m + m->fixed_block + m->fixed_offset[i], cols[i],
				m->fixed_stride[i] * rows);
	}
	if (reuse) {
		v->db.first_tablet[table_id] = (unsigned)tid;
	} else {
		last->last_tablet = 0;
		last->next = (unsigned)tid;
		virg_tablet_unlock(v, last->id);
		virg_tablet_unlock(v, first->id);
		v->db.table_tablets[table_id]++;
	}
	v->db.last_tablet[table_id] = (unsigned)tid;
	v->db.write_cursor[table_id] = (unsigned)tid;
	virg_tablet_unlock(v, (unsigned)tid);
	return VIRG_SUCCESS;
}

/* reference: src/table/loadtomem.c:16.  Everything is resident already. */
int virg_table_loadmem(virginian *v, unsigned table_id)
{
	(void)v;
	(void)table_id;
	return VIRG_SUCCESS;
}

/* reference: src/table/getid.c:18 */
int virg_table_getid(virginian *v, const char *name, unsigned *id)
{
	for (unsigned i = 0; i < VIRG_MAX_TABLES; i++)
		if (v->db.table_status[i] > 0 && strcmp(name, v->db.tables[i]) == 0) {
			id[0] = i;
			return VIRG_SUCCESS;
		}
	return VIRG_FAIL;
}

/* reference: src/table/getcolumn.c:15 */
int virg_table_getcolumn(virginian *v, unsigned tid, const char *name, unsigned *id)
{
	virg_tablet_meta *tab;
	VIRG_CHECK(virg_db_load(v, v->db.first_tablet[tid], &tab) == VIRG_FAIL, "no tablet");
	int found = VIRG_FAIL;
	for (unsigned i = 0; i < tab->fixed_columns; i++)
		if (strcmp(tab->fixed_name[i], name) == 0) {
			id[0] = i;
			found = VIRG_SUCCESS;
			break;
		}
	virg_tablet_unlock(v, tab->id);
	return found;
}

/* reference: src/table/getcolumntype.c:16 */
int virg_table_getcolumntype(virginian *v, unsigned tid, unsigned cid, virg_t *type)
{
	virg_tablet_meta *tab;
	VIRG_CHECK(virg_db_load(v, v->db.first_tablet[tid], &tab) == VIRG_FAIL, "no tablet");
	type[0] = tab->fixed_type[cid];
	virg_tablet_unlock(v, tab->id);
	return VIRG_SUCCESS;
}

/* reference: src/table/getkeytype.c:16 */
int virg_table_getkeytype(virginian *v, unsigned tid, virg_t *type)
{
	virg_tablet_meta *tab;
	VIRG_CHECK(virg_db_load(v, v->db.first_tablet[tid], &tab) == VIRG_FAIL, "no tablet");
	type[0] = tab->key_type;
	virg_tablet_unlock(v, tab->id);
	return VIRG_SUCCESS;
}

/* reference: src/table/numrows.c:16 */
int virg_table_numrows(virginian *v, unsigned id, unsigned *rows)
{
	virg_tablet_meta *tab;
	VIRG_CHECK(virg_db_load(v, v->db.first_tablet[id], &tab) == VIRG_FAIL, "no tablet");
	unsigned n = tab->rows;
	while (!tab->last_tablet) {
		virg_db_loadnext(v, &tab);
		n += tab->rows;
	}
	virg_tablet_unlock(v, tab->id);
	rows[0] = n;
	return VIRG_SUCCESS;
}
