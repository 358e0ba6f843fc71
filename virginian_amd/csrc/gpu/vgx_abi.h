/*
 * Arguments of the specialised kernels (jit.hip).  Plain C over fixed-width
 * integers: the host library and the kernel sources compiled at run time
 * both include this text (the build embeds it, with vg_prog.h, into the
 * library), so both sides see one layout.
 */
#ifndef VGX_ABI_H
#define VGX_ABI_H

#ifndef __HIPCC_RTC__
#include <stdint.h>
#include "vg_prog.h"
#endif

/* a tile of a specialised launch: tablet (index into the descriptors) and
 * its first row in that tablet */
typedef struct {
	uint32_t tab;
	uint32_t row0;
} vgx_tile;

/* kernel arguments; the kernel's own struct has this layout */
typedef struct {
	const vg_tabdesc *tabs;
	const vgx_tile *tiles;
	uint32_t num_tiles;
	uint32_t nchunks;
	uint64_t *state;           /* [0] chunk counter, [1] launch epoch */
	uint64_t *status;          /* one word per chunk */
	vg_ctrl *ctrl;
	uint64_t capacity;
	uint32_t spin_limit;
	uint32_t pad;
	void *out[VG_MAX_OUT];
	uint64_t imm[VG_MAX_UOPS][2];
	uint64_t oimm[VG_MAX_OUT];
} vgx_launch_args;

/* A batch of queries run by one persistent specialised launch: work items
 * are (query, chunk) in a list where every query's chunks are in order;
 * workgroups take them from a counter, so a chunk only ever waits for
 * chunks taken before it. */
typedef struct {
	uint32_t q;                  /* query */
	uint32_t c;                  /* chunk of the query */
} vgxb_item;

typedef struct {
	const vgx_launch_args *qa;   /* per query: its launch (status words, control block, outputs) */
	const uint32_t *fn;          /* per query: its program's index in the module */
	const vgxb_item *items;
	uint32_t nitems;
	uint32_t nquery;
	uint64_t *state;             /* [0] item counter, [1] launch epoch, [2] device clock at the start */
	uint32_t *done;              /* per query: chunks finished in this run */
	vg_bresult *hres;            /* per query: its outcome, in pinned host memory */
} vgxb_args;

#endif
