/*
 * Internal declarations of the host library.
 *
 * Slot manager.  The reference keeps VIRG_MEM_TABLETS fixed 8 MB pinned
 * slots and pages tablets to a database file when they run out
 * (reference: src/db/findslot.c:19-63, src/init.c:53-72).  This library is
 * memory resident: slot i < VIRG_MEM_TABLETS lives in the virginian struct
 * exactly as in the reference, further slots live in an extension table
 * reached through v->gpu_slots (a pointer the reference reserves for the
 * GPU layer, which this library owns).  Slot buffers are allocated on first
 * use and grow to the size a tablet asks for, so one slot can hold a
 * 1M-row tablet.  Nothing is ever written to disk.
 */
#ifndef VIRG_INTERNAL_H
#define VIRG_INTERNAL_H

#include "virginian.h"
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vgd_device vgd_device;   /* device layer state (virg_gpu.c) */

typedef struct {
	unsigned          cap;      /* allocated entries */
	unsigned         *ids;
	int              *status;
	virg_tablet_meta **ptrs;
	size_t           *bytes;    /* allocated size of each extension slot */
} virg_extslots;

/* state hung off v->gpu_slots */
typedef struct {
	unsigned       magic;
	size_t         base_bytes[VIRG_MEM_TABLETS];  /* allocated size of slot i */
	virg_extslots  ext;
	vgd_device    *dev;       /* NULL when no GPU is usable */
	int            pinned;    /* slot memory comes from hipHostMalloc */
	unsigned long long pinned_bytes;    /* slot memory allocated pinned ... */
	unsigned long long pageable_bytes;  /* ... and pageable (pinning failed) */
	unsigned long long pin_limit;       /* pin no slot memory beyond this (0: no limit) */
	/* slot pager: with a database file open, at most resident_limit tablets
	 * stay in host memory (0: no limit); beyond it an unlocked one is
	 * written to the file and its slot reused */
	unsigned       resident_limit;
	unsigned long long evictions;       /* tablets written out to make room */
	unsigned long long file_loads;      /* tablets read back from the file */
} virg_state;

#define VIRG_STATE_MAGIC 0x56474d49u

virg_state *virg_state_of(virginian *v);

/* total number of slots (built-in plus extension) */
unsigned virg_nslots(virginian *v);
unsigned *virg_slot_id(virginian *v, unsigned i);
int *virg_slot_status(virginian *v, unsigned i);
virg_tablet_meta **virg_slot_ptr(virginian *v, unsigned i);
size_t *virg_slot_bytes(virginian *v, unsigned i);

/* find a free slot able to hold `bytes`, mark it used and locked (status 2).
 * Caller holds v->slot_lock. */
int virg_findslot_sized(virginian *v, unsigned *slot, size_t bytes);
/* allocate a tablet slot of at least `bytes` for tablet `id` (locked) */
int virg_db_alloc_sized(virginian *v, virg_tablet_meta **meta, int id, size_t bytes);
/* slot holding tablet id, or -1 */
int virg_slot_of(virginian *v, unsigned id);
/* the disk half of virg_db_load (dbfile.c); the caller holds slot_lock */
int virg_dbfile_load(virginian *v, unsigned tablet_id, virg_tablet_meta **tab);

/* host memory for slots: pinned when a device is present */
void *virg_host_alloc(virginian *v, size_t bytes);
void virg_host_free(virginian *v, void *p);

/* Bulk tablet builder: append a tablet able to hold `capacity` rows to
 * table `table_id` and fill it with `rows` rows from column-major source
 * arrays (key array of key_stride bytes per row, one array per fixed
 * column).  Layout is the reference one (tablet/create.c:52-60,
 * tablet/addcolumn.c:36-37, tablet/addrows.c:39-73 applied to `capacity`
 * rows). */
int virg_table_appendtablet(virginian *v, unsigned table_id, unsigned capacity,
	unsigned rows, const void *key, const void *const *cols);
/* (key and cols NULL: the rows are left for the caller to fill in place) */

/* a result tablet laid out for `rows` rows (vm.c) */
int virg_vm_allocresult_rows(virginian *v, virg_vm *vm, virg_tablet_meta **meta,
	virg_tablet_meta *template_, unsigned rows);

/* the standard synthetic schema (gen.c), no rows yet */
int virg_gen_schema(virginian *v, const char *name, unsigned *table_id);

/* device layer entry points used by the host library */
int vgd_open(virginian *v);            /* 1 if a device is usable */
void vgd_close(virginian *v);
void vgd_forget_tablet(virginian *v, unsigned id);
void *vgd_host_alloc(size_t bytes);    /* pinned, NULL on failure */
void vgd_host_free(void *p);

/* SQL */
int virg_sql_compile(virginian *v, const char *querystr, virg_vm *vm, char *err, size_t errlen);

#ifdef __cplusplus
}
#endif
#endif
