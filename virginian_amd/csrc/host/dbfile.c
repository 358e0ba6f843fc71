/*
 * Database files in the reference's on-disk format, so that a database
 * written by either implementation opens in the other.
 *
 * Layout (reference: src/db/create.c:20-60, open.c:19-50, write.c:50-200,
 * load.c:22-98, close.c:20-58):
 *
 *   offset 0                          the virg_db struct (its tablet_info
 *                                     pointer means nothing on disk)
 *   offset sizeof(virg_db)            alloced_tablets virg_tablet_info entries
 *   block_size + i * VIRG_TABLET_SIZE the tablet entry i describes: tablet
 *                                     ->size bytes, meta block then data
 *
 * Entry i always describes disk slot i.  When every entry is in use the
 * entry table grows by VIRG_TABLET_INFO_INCREMENT; if it would then reach
 * the first tablet, that tablet moves behind the last disk slot, every
 * other slot's index drops by one and block_size grows by one slot, which
 * leaves each tablet where it was on disk.
 *
 * Differences from the reference: once more tablets than the resident
 * limit are in memory, the slot pager (db.c) writes an unlocked one out to
 * make room, as the reference does when its 64 slots are full; tablets are
 * also written when their slot is cleared or the database closes,
 * whatever their lock count; a tablet
 * larger than a disk slot (this library's tablets may hold more rows than
 * the reference's 8 MB) cannot be stored in the format and is refused with
 * a message; a database file cannot be opened once tablets are in memory.
 */
#include "internal.h"
#include <errno.h>
#include <fcntl.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static int pread_all(int fd, void *buf, size_t n, off_t off)
{
	char *p = (char *)buf;
	while (n) {
		ssize_t r = pread(fd, p, n, off);
		if (r < 0 && errno == EINTR)
			continue;
		if (r <= 0)
			return VIRG_FAIL;
		p += r;
		n -= (size_t)r;
		off += r;
	}
	return VIRG_SUCCESS;
}

static int pwrite_all(int fd, const void *buf, size_t n, off_t off)
{
	const char *p = (const char *)buf;
	while (n) {
		ssize_t w = pwrite(fd, p, n, off);
		if (w < 0 && errno == EINTR)
			continue;
		if (w <= 0)
			return VIRG_FAIL;
		p += w;
		n -= (size_t)w;
		off += w;
	}
	return VIRG_SUCCESS;
}

static off_t slot_offset(const virg_db *db, unsigned disk_slot)
{
	return (off_t)db->block_size + (off_t)disk_slot * VIRG_TABLET_SIZE;
}

/* an empty database in memory, with no file */
static void db_reset(virginian *v)
{
	virg_db *db = &v->db;
	free(db->tablet_info);
	db->num_tablets = 0;
	db->tablet_id_counter = 0;
	for (unsigned i = 0; i < VIRG_MAX_TABLES; i++) {
		db->tables[i][0] = '\0';
		db->table_status[i] = 0;
		db->table_tablets[i] = 0;
		db->write_cursor[i] = 0;
		db->first_tablet[i] = 0;
		db->last_tablet[i] = 0;
	}
	db->alloced_tablets = 0;
	db->tablet_info = NULL;
	db->block_size = sizeof(virg_db);
	v->dbfd = -1;
}

/* reference: src/db/create.c:20.  With file == NULL the database lives in
 * memory only (what virg_init sets up). */
int virg_db_create(virginian *v, const char *file)
{
	VIRG_CHECK(v->dbfd != -1, "Database already open");
	db_reset(v);
	if (file == NULL)
		return VIRG_SUCCESS;
	virg_db *db = &v->db;
	size_t size = VIRG_TABLET_INFO_SIZE * sizeof(virg_tablet_info);
	virg_tablet_info *info = (virg_tablet_info *)calloc(VIRG_TABLET_INFO_SIZE, sizeof(virg_tablet_info));
	VIRG_CHECK(info == NULL, "Out of memory");
	for (unsigned i = 0; i < VIRG_TABLET_INFO_SIZE; i++)
		info[i].disk_slot = i;
	int fd = open(file, O_RDWR | O_CREAT | O_EXCL, 0644);
	if (fd < 0) {
		free(info);
		VIRG_CHECK(1, "Problem creating file");
	}
	db->tablet_info = info;
	db->alloced_tablets = VIRG_TABLET_INFO_SIZE;
	db->block_size = sizeof(virg_db) + size;
	v->dbfd = fd;
	return VIRG_SUCCESS;
}

/* reference: src/db/open.c:19 */
int virg_db_open(virginian *v, const char *file)
{
	VIRG_CHECK(v->dbfd != -1, "Database already open");
	VIRG_CHECK(v->tablet_slots_taken != 0, "Tablets are already in memory: open the database file first");
	int fd = open(file, O_RDWR);
	VIRG_CHECK(fd < 0, "Problem opening database file");
	virg_db db;
	int ok = pread_all(fd, &db, sizeof db, 0) == VIRG_SUCCESS &&
		db.alloced_tablets > 0 && db.alloced_tablets <= (1u << 26) &&
		db.block_size >= sizeof(virg_db) + (size_t)db.alloced_tablets * sizeof(virg_tablet_info) &&
		db.num_tablets <= VIRG_MAX_TABLES;
	if (!ok) {
		close(fd);
		VIRG_CHECK(1, "Corrupt database file");
	}
	size_t size = (size_t)db.alloced_tablets * sizeof(virg_tablet_info);
	db.tablet_info = (virg_tablet_info *)malloc(size);
	if (db.tablet_info == NULL || pread_all(fd, db.tablet_info, size, sizeof(virg_db)) == VIRG_FAIL) {
		free(db.tablet_info);
		close(fd);
		VIRG_CHECK(1, "Problem reading tablet info");
	}
	free(v->db.tablet_info);
	v->db = db;
	v->dbfd = fd;
	return VIRG_SUCCESS;
}

/* Grow the entry table (reference: src/db/write.c:66-175); *first_free is
 * set to the first new entry. */
static int grow_info(virginian *v, unsigned *first_free)
{
	virg_db *db = &v->db;
	const unsigned old = db->alloced_tablets, grown = old + VIRG_TABLET_INFO_INCREMENT;
	const size_t size = (size_t)grown * sizeof(virg_tablet_info);
	virg_tablet_info *info = (virg_tablet_info *)calloc(grown, sizeof(virg_tablet_info));
	VIRG_CHECK(info == NULL, "Out of memory");
	const unsigned n = virg_nslots(v);
	if (db->block_size >= size + sizeof(virg_db)) {
		/* the larger table still ends before the first tablet */
		memcpy(info, db->tablet_info, (size_t)old * sizeof(virg_tablet_info));
		for (unsigned s = 0; s < n; s++) {
			virg_tablet_meta *t = *virg_slot_ptr(v, s);
			if (*virg_slot_status(v, s) != 0 && t->info != NULL)
				t->info = &info[t->info->disk_slot];
		}
	} else {
		/* the tablet in disk slot 0 moves behind the last slot */
		memcpy(info, db->tablet_info + 1, (size_t)(old - 1) * sizeof(virg_tablet_info));
		info[old - 1] = db->tablet_info[0];
		for (unsigned i = 0; i < old; i++)
			info[i].disk_slot = i;
		int first_resident = 0;
		for (unsigned s = 0; s < n; s++) {
			virg_tablet_meta *t = *virg_slot_ptr(v, s);
			if (*virg_slot_status(v, s) == 0 || t->info == NULL)
				continue;
			if (t->info->disk_slot == 0) {
				t->info = &info[old - 1];
				first_resident = 1;
			} else {
				t->info = &info[t->info->disk_slot - 1];
			}
		}
		if (!first_resident && db->tablet_info[0].used) {
			/* not in memory: copy its disk image to the new place */
			void *buf = malloc(VIRG_TABLET_SIZE);
			int ok = buf != NULL &&
				pread_all(v->dbfd, buf, VIRG_TABLET_SIZE, slot_offset(db, 0)) == VIRG_SUCCESS &&
				pwrite_all(v->dbfd, buf, VIRG_TABLET_SIZE, slot_offset(db, old)) == VIRG_SUCCESS;
			free(buf);
			if (!ok) {
				free(info);
				VIRG_CHECK(1, "Problem moving a tablet in the database file");
			}
		}
		db->block_size += VIRG_TABLET_SIZE;
	}
	for (unsigned i = old; i < grown; i++) {
		info[i].used = 0;
		info[i].id = 0;
		info[i].disk_slot = i;
	}
	free(db->tablet_info);
	db->tablet_info = info;
	db->alloced_tablets = grown;
	*first_free = old;
	return VIRG_SUCCESS;
}

/* reference: src/db/write.c:50.  The tablet keeps the entry it was given,
 * so writing it again overwrites the same disk slot. */
int virg_db_write(virginian *v, unsigned slot)
{
	VIRG_CHECK(v->dbfd < 0, "No database file is open");
	VIRG_CHECK(slot >= virg_nslots(v) || *virg_slot_status(v, slot) == 0, "Writing an empty tablet slot");
	virg_tablet_meta *tab = *virg_slot_ptr(v, slot);
	VIRG_CHECK(tab->size > VIRG_TABLET_SIZE,
		"Tablet is larger than a database file's disk slot (VIRG_TABLET_SIZE)");
	virg_db *db = &v->db;
	if (tab->info == NULL) {
		unsigned i;
		for (i = 0; i < db->alloced_tablets; i++)
			if (db->tablet_info[i].used == 0)
				break;
		if (i == db->alloced_tablets)
			VIRG_CHECK(grow_info(v, &i) == VIRG_FAIL, "Could not grow the tablet info block");
		db->tablet_info[i].used = 1;
		db->tablet_info[i].id = tab->id;
		db->tablet_info[i].disk_slot = i;
		tab->info = &db->tablet_info[i];
	}
	VIRG_CHECK(pwrite_all(v->dbfd, tab, tab->size, slot_offset(db, tab->info->disk_slot)) == VIRG_FAIL,
		"Failed to write tablet");
	return VIRG_SUCCESS;
}

/* reference: src/db/clear.c:17.  Writes the tablet and frees its slot. */
int virg_db_clear(virginian *v, unsigned slot)
{
	VIRG_CHECK(slot >= virg_nslots(v), "No such tablet slot");
	if (*virg_slot_status(v, slot) == 0)
		return VIRG_SUCCESS;
	VIRG_CHECK(v->dbfd < 0, "No database file is open: clearing the slot would lose the tablet");
	VIRG_CHECK(virg_db_write(v, slot) == VIRG_FAIL, "Could not write the tablet");
	*virg_slot_status(v, slot) = 0;
	v->tablet_slots_taken--;
	vgd_forget_tablet(v, *virg_slot_id(v, slot));
	return VIRG_SUCCESS;
}

/* reference: src/db/close.c:20.  Every tablet in memory is written, then
 * the meta information; the instance is left with an empty database in
 * memory. */
int virg_db_close(virginian *v)
{
	if (v->dbfd == -1)
		return VIRG_SUCCESS;
	int ok = VIRG_SUCCESS;
	const unsigned n = virg_nslots(v);
	for (unsigned s = 0; s < n; s++)
		if (*virg_slot_status(v, s) != 0 && virg_db_clear(v, s) == VIRG_FAIL)
			ok = VIRG_FAIL;
	virg_db *db = &v->db;
	if (pwrite_all(v->dbfd, db, sizeof(virg_db), 0) == VIRG_FAIL ||
		pwrite_all(v->dbfd, db->tablet_info, (size_t)db->alloced_tablets * sizeof(virg_tablet_info),
			sizeof(virg_db)) == VIRG_FAIL) {
		fprintf(stderr, "Problem writing db meta info\n");
		ok = VIRG_FAIL;
	}
	if (close(v->dbfd) != 0) {
		fprintf(stderr, "Problem closing file\n");
		ok = VIRG_FAIL;
	}
	v->dbfd = -1;
	db_reset(v);
	return ok;
}

/* The disk half of virg_db_load (reference: src/db/load.c:44-93); the
 * caller holds slot_lock.  The tablet comes back resident with one lock. */
int virg_dbfile_load(virginian *v, unsigned tablet_id, virg_tablet_meta **tab)
{
	virg_db *db = &v->db;
	unsigned i;
	for (i = 0; i < db->alloced_tablets; i++)
		if (db->tablet_info[i].used == 1 && db->tablet_info[i].id == tablet_id)
			break;
	VIRG_CHECK(i == db->alloced_tablets, "Could not find tablet id");
	const off_t off = slot_offset(db, i);
	virg_tablet_meta m;
	VIRG_CHECK(pread_all(v->dbfd, &m, sizeof m, off) == VIRG_FAIL, "Failed to get tablet meta data");
	VIRG_CHECK(m.size < sizeof m || m.size > VIRG_TABLET_SIZE || m.id != tablet_id,
		"Corrupt tablet in database file");
	unsigned slot;
	VIRG_CHECK(virg_findslot_sized(v, &slot, VIRG_TABLET_SIZE) == VIRG_FAIL, "No tablet slot");
	virg_tablet_meta *t = *virg_slot_ptr(v, slot);
	if (pread_all(v->dbfd, t, m.size, off) == VIRG_FAIL) {
		*virg_slot_status(v, slot) = 0;
		v->tablet_slots_taken--;
		VIRG_CHECK(1, "Failed to get tablet data");
	}
	t->info = &db->tablet_info[i];
	*virg_slot_id(v, slot) = tablet_id;
	virg_state_of(v)->file_loads++;
	if (tab != NULL)
		tab[0] = t;
	return VIRG_SUCCESS;
}
