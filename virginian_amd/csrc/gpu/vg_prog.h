/*
 * Decoded op-program: the form in which the data-parallel section of a
 * virg_vm program reaches the gfx950 kernels.
 *
 * The host decoder (decode.cpp) resolves everything that is uniform across
 * rows before launch: register types (the reference keeps one type per
 * register for a whole row block, src/vm/virginia.c:449 and :144-145),
 * constants (Integer/Float registers become immediates), column loads
 * (Column/Rowid copy a whole block in the reference, virginia.c:445-448, so
 * a column register is simply the loaded column), jump targets (remapped to
 * micro-op indices) and physical register slots.  The kernel then walks the
 * micro-ops with wave-uniform dispatch: every field below is read with
 * scalar loads.
 *
 * Physical register file: up to VG_NSLOT 32-bit slots per row, held in
 * LDS as [slot][row of the tile].  Loaded columns take the first slots
 * (8-byte types two: low and high word), computed values follow; the
 * decoder reuses a computed value's slot once its last reader has run.
 */
#ifndef VG_PROG_H
#define VG_PROG_H

#ifndef __HIPCC_RTC__
#include <stdint.h>
#endif

#define VG_MAX_UOPS      48
#define VG_MAX_COLS      17      /* distinct column sources: 16 fixed + key */
#define VG_MAX_OUT       16
#define VG_NSLOT         48
#define VG_KEY_SRC       16      /* column source index of the key column */

/* micro-op kinds */
enum {
	VG_K_LT = 0, VG_K_LE, VG_K_GT, VG_K_GE, VG_K_EQ, VG_K_NE, VG_K_AND, VG_K_OR,
	VG_K_NOT,        /* branch if a == 0 */
	VG_K_INVALID,    /* active rows become invalid */
	VG_K_ADD, VG_K_SUB, VG_K_MUL, VG_K_DIV,
	VG_K_MOV,        /* dst = a (typed copy) */
	VG_K_CAST,       /* dst = (type b) a, in place on the reference's register */
	VG_K_COUNT
};

/* flags */
#define VG_F_AIMM    0x01   /* operand a is the immediate ia */
#define VG_F_BIMM    0x02   /* operand b is the immediate ib */
#define VG_F_MASKED  0x04   /* write only rows active at this micro-op */
#define VG_F_KEEP    0x08   /* branch keeps validity (p4 != 0) */

/* value types: the virg_t numbering */
enum { VG_T_INT = 0, VG_T_INT64 = 1, VG_T_FLOAT = 2, VG_T_DOUBLE = 3, VG_T_CHAR = 4 };

typedef struct {
	uint8_t  kind;
	uint8_t  type;     /* operand type; for CAST the source type */
	uint8_t  flags;
	uint8_t  target;   /* branch: micro-op index where taken rows resume */
	uint8_t  dst, dst_hi;
	uint8_t  a, a_hi;
	uint8_t  b, b_hi;  /* for CAST: b = destination type */
	uint16_t src_op;   /* op index in the original program (diagnostics) */
	uint8_t  wslot;    /* branch with VG_F_KEEP: wait slot of the rows it takes */
	uint8_t  join;     /* bit j: rows waiting in slot j become active here */
	uint16_t pad;
	uint64_t ia, ib;   /* immediates (bit patterns, 32-bit types in the low word) */
} vg_uop;              /* 32 bytes */

/* Rows that take a forward branch which keeps their validity wait until
 * the branch target; at most VG_NWAIT different targets can be pending at
 * once (the decoder assigns the slots). */
#define VG_NWAIT 4

/* column loads */
enum { VG_PHASE_FILTER = 0, VG_PHASE_PROJECT = 1, VG_PHASE_EMIT = 2 };

typedef struct {
	uint8_t src;       /* 0..15 fixed column, VG_KEY_SRC = key */
	uint8_t width;     /* 1, 4 or 8 bytes */
	uint8_t phase;
	uint8_t type;
	uint8_t slot, slot_hi;
	uint8_t wload;     /* loaded again by the write pass */
	uint8_t pad;
} vg_col;              /* 8 bytes */

/* result columns */
enum { VG_OUT_SLOT = 0, VG_OUT_COLUMN = 1, VG_OUT_CONST = 2 };

typedef struct {
	uint8_t  kind;
	uint8_t  width;    /* bytes written per row */
	uint8_t  slot, slot_hi;
	uint8_t  col;      /* VG_OUT_COLUMN: index into vg_prog.col */
	uint8_t  type;
	uint8_t  pad[2];
	uint64_t imm;      /* VG_OUT_CONST value (bit pattern) */
} vg_out;              /* 16 bytes */

typedef struct {
	uint32_t nuop;      /* micro-ops in total */
	uint32_t nfilter;   /* micro-ops [0, nfilter) can change validity */
	uint32_t ncol;      /* entries in col[] (loaded ones first) */
	uint32_t nload;     /* entries of col[] that occupy slots */
	uint32_t nout;
	uint32_t nslot;     /* highest slot used + 1 */
	uint32_t has_project_loads;
	uint32_t wfull;     /* the write pass reruns every micro-op (masked writes) */
	uint64_t wneed;     /* micro-ops the write pass runs (bit k = micro-op k) */
	uint32_t nslot_count; /* slots [0, nslot_count) hold everything the count pass uses */
	uint32_t wsafe;     /* no output reads a loaded column's slot: the write pass
	                     * may load the next tile while writing this one */
	vg_col   col[VG_MAX_COLS + 1];
	vg_out   out[VG_MAX_OUT];
	vg_uop   uop[VG_MAX_UOPS];
} vg_prog;

/* per-tablet device descriptor: base address of every column of one
 * device-resident tablet, its row count and its first tile */
typedef struct {
	uint64_t col[VG_MAX_COLS];   /* device address of column (or key) row 0 */
	uint32_t rows;
	uint32_t first_tile;
	uint32_t aligned;            /* every column base is 16-byte aligned */
	uint32_t pad;
} vg_tabdesc;                    /* 152 bytes */

/* per-launch control block in device memory */
typedef struct {
	uint64_t total;          /* rows selected (written by the last workgroup) */
	uint32_t overflow;       /* output capacity was exceeded */
	uint32_t pad[5];
} vg_ctrl;

/* A query's outcome of a run, stored by the kernel that finishes the
 * query into pinned host memory (a batch's reset, or the host before a
 * lone run, clears it): the host reads it where it lies, with no copy that
 * would queue behind kernels still holding the compute units.  `done` is
 * stored last (system scope, release). */
typedef struct {
	uint64_t total;          /* rows selected */
	uint64_t t_done;         /* device clock when its last item finished */
	uint64_t t_start;        /* device clock when the run started */
	uint32_t flags;          /* overflow bits */
	uint32_t done;           /* 1 once the fields above are final */
} vg_bresult;

/* kernel arguments besides the program.  Count-pass workgroup g owns
 * tiles [g * tiles_per_wg, (g + 1) * tiles_per_wg), write-pass workgroup g
 * tiles [g * wtiles_per_wg, (g + 1) * wtiles_per_wg); each pass has the
 * grid its own occupancy allows. */
typedef struct {
	const vg_tabdesc *tabs;
	const uint32_t *tile_tab;
	uint32_t num_tiles;
	uint32_t tiles_per_wg;
	uint32_t wtiles_per_wg;
	uint32_t pad;
	uint64_t capacity;
	vg_ctrl *ctrl;
	uint8_t *vbits;          /* one validity byte per write-pass lane per tile */
	uint32_t *tcount;        /* rows selected per tile and wave: 4 per tile */
	uint32_t *wgcount;       /* rows selected per count-pass workgroup */
	vg_bresult *hres;        /* NULL, or where a launch after the write pass
	                          * stores the run's outcome (pinned host memory) */
	void *out[VG_MAX_OUT];
} vg_launch;

/* ---- batched execution: persistent launches run many queries ----
 * Work items are (query, tile range).  Every query's tiles are cut into
 * items of at most tiles_per_item tiles; count item k and write item k of
 * a query cover the same tiles.  The count launch works through the count
 * list and the write launch through the write list (any order each; a
 * query's write items are consecutive). */
typedef struct {
	uint32_t query;
	uint32_t t0, t1;         /* tiles [t0, t1) of the query's launch */
	uint32_t k;              /* item index within the query */
} vg_item;

typedef struct {
	uint32_t nitem;          /* items of the query in each pass */
	uint32_t ioff;           /* its first entry of icount / ibase */
	uint32_t soff;           /* its first tile status word (single pass) */
	uint32_t wstart;         /* its first item in its group's write list */
} vg_bqinfo;

/* Single pass: one status word per tile, published by the workgroup that
 * holds the tile: bits 63:62 = 0 not yet, 1 the tile's own count, 2 the
 * count of the query's rows up to and including the tile; bits 61:0 the
 * count.  Cleared before every run. */
#define VG_ST_AGG  (1ull << 62)
#define VG_ST_PRE  (2ull << 62)
#define VG_ST_VAL  ((1ull << 62) - 1)
/* vg_ctrl.overflow bits */
#define VG_OVF_CAPACITY 1u       /* output capacity exceeded */
#define VG_OVF_WAIT     2u       /* a single-pass wait gave up: result incomplete */

/* per-query state, cleared before every run */
typedef struct {
	uint32_t done;           /* count items finished */
	uint32_t flags;          /* the query's overflow bits, when finished */
	uint32_t wdone;          /* write (or single-pass) items finished */
	uint32_t next;           /* single pass: the query's next tile item */
	uint64_t t_done;         /* device clock when its last item finished */
	uint64_t total;          /* rows selected, when finished */
} vg_bquery;

typedef struct {
	uint64_t t_start;        /* device clock when the run started */
} vg_bhead;

typedef struct {
	const vg_prog *progs;    /* one per query */
	const vg_launch *launch; /* one per query */
	const vg_bqinfo *info;
	const vg_item *items;    /* count list */
	const vg_item *witems;   /* write list (the same items, maybe reordered) */
	uint32_t nitems;
	uint32_t nquery;
	vg_bquery *q;
	uint32_t *icount;        /* rows selected by each count item */
	uint64_t *ibase;         /* output row of each item's first selected row */
	vg_bhead *head;          /* this list's head (the reset clears nhead of them) */
	uint32_t nhead;
	uint32_t spin_limit;     /* single pass: polls of a predecessor before giving up */
	uint64_t *status;        /* single pass: tile status words of every query */
	uint64_t nstatus;
	uint64_t *stamp;         /* device clock when each of the group's write items finished */
	const uint32_t *gq;      /* the queries of this launch's group */
	uint32_t ngq;
	uint32_t pad2;
	vg_bresult *hres;        /* per query, in pinned host memory */
} vg_batch_args;

#endif
