/*
 * Process state, slot manager and tablet locks.
 *
 * Follows the reference's slot protocol: a slot's status is 0 when free,
 * 1 when it holds an unlocked tablet and 1+k with k locks
 * (reference: src/db/findslot.c:19-63, src/tablet/lock.c:20-98,
 * src/db/load.c:22-98, src/db/loadnext.c:21-33, src/db/alloc.c:18-42).
 * Differences: eviction happens only with a database file open (without
 * one a tablet has nowhere to go), and then once more than resident_limit
 * tablets (VIRG_MEM_TABLETS unless set with vg_tablet_limit) are in host
 * memory, not when a fixed slot table is full; tablets larger than a disk
 * slot are never chosen; slot buffers are sized per tablet, and slots beyond
 * VIRG_MEM_TABLETS come from an extension table.
 */
#include "internal.h"
#include <stdlib.h>
#include <string.h>

virg_state *virg_state_of(virginian *v)
{
	virg_state *s = (virg_state *)v->gpu_slots;
	return (s && s->magic == VIRG_STATE_MAGIC) ? s : NULL;
}

unsigned virg_nslots(virginian *v)
{
	virg_state *s = virg_state_of(v);
	return VIRG_MEM_TABLETS + (s ? s->ext.cap : 0);
}

unsigned *virg_slot_id(virginian *v, unsigned i)
{
	if (i < VIRG_MEM_TABLETS)
		return &v->tablet_slot_ids[i];
	return &virg_state_of(v)->ext.ids[i - VIRG_MEM_TABLETS];
}

int *virg_slot_status(virginian *v, unsigned i)
{
	if (i < VIRG_MEM_TABLETS)
		return &v->tablet_slot_status[i];
	return &virg_state_of(v)->ext.status[i - VIRG_MEM_TABLETS];
}

virg_tablet_meta **virg_slot_ptr(virginian *v, unsigned i)
{
	if (i < VIRG_MEM_TABLETS)
		return &v->tablet_slots[i];
	return &virg_state_of(v)->ext.ptrs[i - VIRG_MEM_TABLETS];
}

size_t *virg_slot_bytes(virginian *v, unsigned i)
{
	virg_state *s = virg_state_of(v);
	if (i < VIRG_MEM_TABLETS)
		return &s->base_bytes[i];
	return &s->ext.bytes[i - VIRG_MEM_TABLETS];
}

void *virg_host_alloc(virginian *v, size_t bytes)
{
	virg_state *s = virg_state_of(v);
	if (s && s->pinned && (!s->pin_limit || s->pinned_bytes + bytes <= s->pin_limit)) {
		void *p = vgd_host_alloc(bytes);
		if (p) {
			s->pinned_bytes += bytes;
			return p;
		}
	}
	void *p = NULL;
	if (posix_memalign(&p, 4096, bytes) != 0)
		return NULL;
	if (s)
		s->pageable_bytes += bytes;
	return p;
}

void virg_host_free(virginian *v, void *p)
{
	virg_state *s = virg_state_of(v);
	if (!p)
		return;
	if (s && s->pinned)
		vgd_host_free(p);
	else
		free(p);
}

static int ext_grow(virginian *v)
{
	virg_state *s = virg_state_of(v);
	unsigned ncap = s->ext.cap ? s->ext.cap * 2 : 64;
	unsigned *ids = realloc(s->ext.ids, ncap * sizeof(unsigned));
	if (ids) s->ext.ids = ids;
	int *st = realloc(s->ext.status, ncap * sizeof(int));
	if (st) s->ext.status = st;
	virg_tablet_meta **pt = realloc(s->ext.ptrs, ncap * sizeof(void *));
	if (pt) s->ext.ptrs = pt;
	size_t *by = realloc(s->ext.bytes, ncap * sizeof(size_t));
	if (by) s->ext.bytes = by;
	VIRG_CHECK(!ids || !st || !pt || !by, "Out of memory growing the slot table");
	for (unsigned i = s->ext.cap; i < ncap; i++) {
		s->ext.ids[i] = 0;
		s->ext.status[i] = 0;
		s->ext.ptrs[i] = NULL;
		s->ext.bytes[i] = 0;
	}
	s->ext.cap = ncap;
	return VIRG_SUCCESS;
}

/* Make room: the first unlocked tablet from the round-robin position on
 * (reference: src/db/findslot.c:36-57) is written to the database file and
 * its slot freed.  Caller holds v->slot_lock. */
static int evict_one(virginian *v)
{
	virg_state *s = virg_state_of(v);
	const unsigned n = virg_nslots(v);
	for (unsigned k = 0; k < n; k++) {
		const unsigned slot = (v->tablet_slot_counter + k) % n;
		if (*virg_slot_status(v, slot) != 1)
			continue;
		const virg_tablet_meta *t = *virg_slot_ptr(v, slot);
		if (t->size > VIRG_TABLET_SIZE)
			continue;               /* no disk slot holds it */
		v->tablet_slot_counter = (slot + 1) % n;
		VIRG_CHECK(virg_db_write(v, slot) == VIRG_FAIL, "Could not write a tablet out to make room");
		*virg_slot_status(v, slot) = 0;
		v->tablet_slots_taken--;
		s->evictions++;
		return VIRG_SUCCESS;
	}
	VIRG_CHECK(1, "All tablets locked");
	return VIRG_FAIL;
}

int virg_findslot_sized(virginian *v, unsigned *slot_, size_t bytes)
{
	virg_state *s = virg_state_of(v);
	if (v->dbfd >= 0 && s->resident_limit && v->tablet_slots_taken >= s->resident_limit)
		VIRG_CHECK(evict_one(v) == VIRG_FAIL, "No tablet slot");
	unsigned n = virg_nslots(v);
	unsigned slot;
	/* a free slot whose buffer is big enough, else the first free one */
	for (slot = 0; slot < n; slot++)
		if (*virg_slot_status(v, slot) == 0 && *virg_slot_ptr(v, slot) && *virg_slot_bytes(v, slot) >= bytes)
			break;
	if (slot == n)
		for (slot = 0; slot < n; slot++)
			if (*virg_slot_status(v, slot) == 0)
				break;
	if (slot == n) {
		VIRG_CHECK(ext_grow(v) == VIRG_FAIL, "All tablet slots in use");
		slot = n;
	}
	size_t *have = virg_slot_bytes(v, slot);
	virg_tablet_meta **ptr = virg_slot_ptr(v, slot);
	if (*ptr == NULL || *have < bytes) {
		if (*ptr)
			virg_host_free(v, *ptr);
		*ptr = (virg_tablet_meta *)virg_host_alloc(v, bytes);
		VIRG_CHECK(*ptr == NULL, "Could not allocate tablet slot memory");
		*have = bytes;
	}
	v->tablet_slots_taken++;
	*virg_slot_status(v, slot) = 2;
	slot_[0] = slot;
	return VIRG_SUCCESS;
}

/* reference: src/db/findslot.c:19 (memory-resident variant) */
int virg_db_findslot(virginian *v, unsigned *slot_)
{
	return virg_findslot_sized(v, slot_, VIRG_TABLET_SIZE);
}

int virg_db_alloc_sized(virginian *v, virg_tablet_meta **meta, int id, size_t bytes)
{
	unsigned slot;
	pthread_mutex_lock(&v->slot_lock);
	int ok = virg_findslot_sized(v, &slot, bytes);
	if (ok == VIRG_SUCCESS) {
		meta[0] = *virg_slot_ptr(v, slot);
		*virg_slot_id(v, slot) = (unsigned)id;
		memset(meta[0], 0, sizeof(virg_tablet_meta));
		meta[0]->id = (unsigned)id;
	}
	pthread_mutex_unlock(&v->slot_lock);
	return ok;
}

/* reference: src/db/alloc.c:18 */
int virg_db_alloc(virginian *v, virg_tablet_meta **meta, int id)
{
	return virg_db_alloc_sized(v, meta, id, VIRG_TABLET_SIZE);
}

int virg_slot_of(virginian *v, unsigned id)
{
	unsigned n = virg_nslots(v);
	for (unsigned i = 0; i < n; i++)
		if (*virg_slot_status(v, i) != 0 && *virg_slot_id(v, i) == id)
			return (int)i;
	return -1;
}

/* reference: src/db/load.c:22.  Adds a lock to a resident tablet, or reads
 * it from the database file into a slot (with one lock). */
int virg_db_load(virginian *v, unsigned tablet_id, virg_tablet_meta **tab)
{
	pthread_mutex_lock(&v->slot_lock);
	int i = virg_slot_of(v, tablet_id);
	int ok = VIRG_SUCCESS;
	if (i >= 0) {
		if (tab != NULL)
			tab[0] = *virg_slot_ptr(v, (unsigned)i);
		(*virg_slot_status(v, (unsigned)i))++;
	} else if (v->dbfd >= 0) {
		ok = virg_dbfile_load(v, tablet_id, tab);
	} else {
		ok = VIRG_FAIL;
	}
	pthread_mutex_unlock(&v->slot_lock);
	VIRG_CHECK(ok == VIRG_FAIL, "Could not find tablet id");
	return VIRG_SUCCESS;
}

/* reference: src/db/loadnext.c:21 */
int virg_db_loadnext(virginian *v, virg_tablet_meta **tab)
{
	virg_tablet_meta *t = tab[0];
	VIRG_CHECK(virg_db_load(v, t->next, tab) == VIRG_FAIL, "Could not load next tablet");
	virg_tablet_unlock(v, t->id);
	return VIRG_SUCCESS;
}

/* reference: src/tablet/lock.c:20 */
int virg_tablet_lock(virginian *v, unsigned tablet_id)
{
	pthread_mutex_lock(&v->slot_lock);
	int i = virg_slot_of(v, tablet_id);
	if (i >= 0)
		(*virg_slot_status(v, (unsigned)i))++;
	pthread_mutex_unlock(&v->slot_lock);
	VIRG_CHECK(i < 0, "Couldn't find tablet to lock");
	return VIRG_SUCCESS;
}

/* reference: src/tablet/lock.c:67 */
int virg_tablet_unlock(virginian *v, unsigned tablet_id)
{
	pthread_mutex_lock(&v->slot_lock);
	unsigned n = virg_nslots(v);
	unsigned i;
	for (i = 0; i < n; i++)
		if (*virg_slot_status(v, i) > 1 && *virg_slot_id(v, i) == tablet_id)
			break;
	if (i < n)
		(*virg_slot_status(v, i))--;
	pthread_mutex_unlock(&v->slot_lock);
	VIRG_CHECK(i == n, "Couldn't find tablet to unlock");
	return VIRG_SUCCESS;
}

/* reference: src/tablet/remove.c:18.  The slot keeps its buffer for reuse;
 * the tablet's entry in the database file, if any, is freed. */
int virg_tablet_remove(virginian *v, unsigned id)
{
	pthread_mutex_lock(&v->slot_lock);
	int i = virg_slot_of(v, id);
	int found = i >= 0;
	if (found) {
		virg_tablet_meta *t = *virg_slot_ptr(v, (unsigned)i);
		if (t->info != NULL)
			t->info->used = 0;
		*virg_slot_status(v, (unsigned)i) = 0;
		v->tablet_slots_taken--;
	} else {
		virg_db *db = &v->db;
		for (unsigned k = 0; k < db->alloced_tablets && !found; k++)
			if (db->tablet_info[k].used == 1 && db->tablet_info[k].id == id) {
				db->tablet_info[k].used = 0;
				found = 1;
			}
	}
	pthread_mutex_unlock(&v->slot_lock);
	VIRG_CHECK(!found, "Could not find tablet to remove");
	vgd_forget_tablet(v, id);
	return VIRG_SUCCESS;
}

/* reference: src/init.c:19.  Device setup moves to vgd_open; slot memory
 * is allocated on first use instead of 64 x 8 MB up front. */
int virg_init(virginian *v)
{
	memset(v, 0, sizeof(*v));
	v->threads_per_block = VIRG_THREADSPERBLOCK;
	v->multi_threads = VIRG_MULTITHREADS;
	v->dbfd = -1;
	VIRG_CHECK(pthread_mutex_init(&v->slot_lock, NULL) != 0, "Could not init mutex");
	virg_state *s = (virg_state *)calloc(1, sizeof(virg_state));
	VIRG_CHECK(s == NULL, "Out of memory");
	s->magic = VIRG_STATE_MAGIC;
	s->resident_limit = VIRG_MEM_TABLETS;
	v->gpu_slots = s;
	virg_db_create(v, NULL);
	s->pinned = vgd_open(v);
	return VIRG_SUCCESS;
}

/* reference: src/close.c:20 (an open database file is closed first) */
int virg_close(virginian *v)
{
	virg_state *s = virg_state_of(v);
	VIRG_CHECK(s == NULL, "virginian state not initialised");
	int ok = virg_db_close(v);
	unsigned n = virg_nslots(v);
	for (unsigned i = 0; i < n; i++) {
		virg_tablet_meta **p = virg_slot_ptr(v, i);
		if (*p)
			virg_host_free(v, *p);
		*p = NULL;
	}
	vgd_close(v);
	free(s->ext.ids);
	free(s->ext.status);
	free(s->ext.ptrs);
	free(s->ext.bytes);
	s->magic = 0;
	free(s);
	v->gpu_slots = NULL;
	pthread_mutex_destroy(&v->slot_lock);
	return ok;
}

/* Slot memory allocated since virg_init: pinned, and pageable (no device,
 * or pinning failed). */
int vg_host_memory(virginian *v, unsigned long long *pinned, unsigned long long *pageable)
{
	virg_state *s = virg_state_of(v);
	VIRG_CHECK(s == NULL, "database not open");
	*pinned = s->pinned_bytes;
	*pageable = s->pageable_bytes;
	return VIRG_SUCCESS;
}

/* Pin at most `bytes` of tablet slot memory (0: as much as the system
 * allows); later slots are pageable.  Leaves pinned memory for staging and
 * result buffers when a table is larger than what the system will pin. */
int vg_host_pin_limit(virginian *v, unsigned long long bytes)
{
	virg_state *s = virg_state_of(v);
	VIRG_CHECK(s == NULL, "database not open");
	s->pin_limit = bytes;
	return VIRG_SUCCESS;
}

/* At most `tablets` tablets in host memory while a database file is open
 * (0: no limit); beyond it unlocked tablets are written to the file and
 * read back when next loaded.  Returns the previous limit. */
unsigned vg_tablet_limit(virginian *v, unsigned tablets)
{
	virg_state *s = virg_state_of(v);
	if (s == NULL)
		return 0;
	pthread_mutex_lock(&v->slot_lock);
	const unsigned old = s->resident_limit;
	s->resident_limit = tablets;
	pthread_mutex_unlock(&v->slot_lock);
	return old;
}

/* Tablets in host memory, tablets written out to make room and tablets
 * read back from the database file since virg_init; *resident_id (when not
 * NULL) = 1 if tablet `id` is in host memory. */
int vg_pager_stats(virginian *v, unsigned *resident, unsigned long long *evictions,
	unsigned long long *file_loads, unsigned id, int *resident_id)
{
	virg_state *s = virg_state_of(v);
	VIRG_CHECK(s == NULL, "database not open");
	pthread_mutex_lock(&v->slot_lock);
	if (resident)
		*resident = v->tablet_slots_taken;
	if (evictions)
		*evictions = s->evictions;
	if (file_loads)
		*file_loads = s->file_loads;
	if (resident_id)
		*resident_id = virg_slot_of(v, id) >= 0;
	pthread_mutex_unlock(&v->slot_lock);
	return VIRG_SUCCESS;
}
