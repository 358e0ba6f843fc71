/*
 * virginian.h -- public C API of the MI355X opcode-VM query executor.
 *
 * This header keeps the binary interface of the Virginian database API
 * (reference: src/virginian.h): the opcode enumeration, the type tags, and
 * every struct that crosses a function boundary have the same members, in
 * the same order, with the same sizes and alignment, so that code compiled
 * against the reference header links against this library unchanged.
 * The struct sizes are pinned by tests (reference: src/test/virgstate.cc:64-80).
 *
 * What differs from the reference is below the structs: the data-parallel
 * VM runs as hand-written gfx950 kernels (virg_vm_gpu), the host slot
 * manager keeps tablets in pinned host memory and a device-resident tablet
 * cache, and the SQL front-end is a hand-written recursive-descent parser.
 * No CUDA or HIP type appears here; device work is reached only through
 * the functions declared below and in virg_gpu.h.
 */
#ifndef VIRGINIAN_H
#define VIRGINIAN_H

#include <stddef.h>
#include <pthread.h>
#include <sys/time.h>

/* ---- sizes and limits (reference: src/virginian.h:47-115) ------------- */
#define VIRG_KB                     1024
#define VIRG_MB                     1048576
#define VIRG_GB                     1073741824

/* acceptable error in floating point comparisons (the reference GPU VM
 * uses it for float Eq; this library and the reference CPU VM compare
 * exactly) */
#define VIRG_FLOAT_ERROR            0.0001

#define VIRG_TABLET_SIZE            (8 * VIRG_MB)
#define VIRG_TABLET_INITIAL_KEYS    256
#define VIRG_TABLET_KEY_INCREMENT   (2048 * 128)
#define VIRG_TABLET_INITIAL_FIXED   0
#define VIRG_TABLET_INITIAL_VARIABLE 0
#define VIRG_TABLET_MAXED_VARIABLE  (VIRG_TABLET_SIZE / 16)
#define VIRG_RESULT_INITIAL_VARIABLE (512 * VIRG_KB)
#define VIRG_TABLET_INFO_SIZE       16
#define VIRG_TABLET_INFO_INCREMENT  32

#define VIRG_MAX_COLUMNS            16
#define VIRG_MAX_TABLES             16
#define VIRG_MAX_COLUMN_NAME        16
#define VIRG_MAX_TABLE_NAME         32

#define VIRG_MEM_TABLETS            64
#define VIRG_GPU_TABLETS            2
#define VIRG_VM_TABLES              1
#define VIRG_CPU_SIMD               64
#define VIRG_ROW_BUFFER             256
#define VIRG_REGS                   16
#define VIRG_GLOBAL_REGS            16
/* Op-program capacity of virg_vm (kept at the reference value so virg_vm
 * stays binary compatible).  Longer programs go through the op-array entry
 * points of virg_gpu.h, which take (virg_op*, count) directly. */
#define VIRG_OPS                    32

#define VIRG_CUDADEVICE             0
#define VIRG_THREADSPERBLOCK        128
#define VIRG_THREADSPERBLOCK_MASK   0xFFFFFF80
#define VIRG_MULTITHREADS           8

#define VIRG_FAIL                   0
#define VIRG_SUCCESS                1

#define VIRG_MIN(a, b)              ((a) < (b) ? (a) : (b))
#define VIRG_MAX(a, b)              ((a) < (b) ? (b) : (a))
#define VIRG_ISPWR2(x)              (((x) != 0) && (((x) & (~(x) + 1)) == (x)))

#include <stdio.h>
#define VIRG_ERROR(msg) \
	fprintf(stderr, "::: Virginian error %s line %d:: %s\n", __FILE__, __LINE__, (msg))
#define VIRG_CHECK(cond, msg) \
	do { if (cond) { VIRG_ERROR(msg); return VIRG_FAIL; } } while (0)

#ifdef __cplusplus
extern "C" {
#endif

/* ---- opcodes (reference: src/virginian.h:166-193) --------------------- */
typedef enum {
	OP_Table = 0, OP_ResultColumn = 1, OP_Parallel = 2, OP_Finish = 3,
	OP_Column = 4, OP_Rowid = 5, OP_Result = 6, OP_Converge = 7,
	OP_Invalid = 8, OP_Cast = 9, OP_Integer = 10, OP_Float = 11,
	OP_Le = 12, OP_Lt = 13, OP_Ge = 14, OP_Gt = 15, OP_Eq = 16, OP_Neq = 17,
	OP_Add = 18, OP_Sub = 19, OP_Mul = 20, OP_Div = 21,
	OP_And = 22, OP_Or = 23, OP_Not = 24, OP_Nop = 25
} virg_ops;

/* ---- value types (reference: src/virginian.h:197-219) ----------------- */
typedef enum {
	VIRG_INT = 0, VIRG_INT64 = 1, VIRG_FLOAT = 2, VIRG_DOUBLE = 3,
	VIRG_CHAR = 4, VIRG_STRING = 5, VIRG_NULL = 6
} virg_t;

virg_t virg_generalizetype(virg_t t1, virg_t t2);

static const size_t virg_sizes[7] = {
	sizeof(int), sizeof(long long int), sizeof(float), sizeof(double),
	sizeof(char), sizeof(char) * 4, 0
};

/* ---- tablets (reference: src/virginian.h:229-302) --------------------- */
typedef struct {
	unsigned id;
	int      used;
	unsigned disk_slot;
} virg_tablet_info;

/* Meta block at the start of every tablet.  Cell (row r, fixed column j)
 * lives at  (char*)tab + fixed_block + fixed_offset[j] + fixed_stride[j]*r,
 * and the key of row r at (char*)tab + key_block + key_stride*r. */
typedef struct {
	unsigned  rows;
	virg_t    key_type;
	size_t    key_stride;
	size_t    key_pointer_stride;
	unsigned  id;
	unsigned  next;
	int       last_tablet;
	int       in_table;
	unsigned  table_id;
	size_t    key_block;
	size_t    key_pointers_block;
	size_t    fixed_block;
	size_t    variable_block;
	size_t    size;
	unsigned  possible_rows;
	size_t    row_stride;
	unsigned  fixed_columns;
	char      fixed_name[VIRG_MAX_COLUMNS][VIRG_MAX_COLUMN_NAME];
	virg_t    fixed_type[VIRG_MAX_COLUMNS];
	size_t    fixed_stride[VIRG_MAX_COLUMNS];
	size_t    fixed_offset[VIRG_MAX_COLUMNS];
	virg_tablet_info *info;
} __attribute__((aligned(64))) virg_tablet_meta;

/* ---- op-program (reference: src/virginian.h:308-332) ------------------ */
typedef union {
	int           i;
	float         f;
	long long int li;
	double        d;
	char          c;
	char         *s;
} __attribute__((aligned(8))) virg_var;

typedef struct {
	int      op;
	int      p1;
	int      p2;
	int      p3;
	virg_var p4;
} virg_op;

typedef struct virg_result_node_ {
	unsigned id;
	struct virg_result_node_ *next;
} virg_result_node;

/* reference: src/virginian.h:356-377 */
typedef struct {
	unsigned          pc;
	virg_op           stmt[VIRG_OPS];
	unsigned          num_ops;
	virg_var          global_reg[VIRG_GLOBAL_REGS];
	virg_t            type[VIRG_GLOBAL_REGS];
	unsigned          table[VIRG_VM_TABLES];
	unsigned          num_tables;
	virg_result_node *head_result;
	virg_result_node *tail_result;
	float             timing1, timing2, timing3;
} virg_vm;

/* ---- database state (reference: src/virginian.h:440-504) -------------- */
typedef struct {
	unsigned num_tablets;
	unsigned alloced_tablets;
	unsigned tablet_id_counter;
	size_t   block_size;
	char     tables[VIRG_MAX_TABLES][VIRG_MAX_TABLE_NAME];
	unsigned first_tablet[VIRG_MAX_TABLES];
	unsigned last_tablet[VIRG_MAX_TABLES];
	unsigned write_cursor[VIRG_MAX_TABLES];
	unsigned table_tablets[VIRG_MAX_TABLES];
	int      table_status[VIRG_MAX_TABLES];
	virg_tablet_info *tablet_info;
} virg_db;

typedef struct {
	virg_db           db;
	unsigned          tablet_slot_ids[VIRG_MEM_TABLETS];
	int               tablet_slot_status[VIRG_MEM_TABLETS];
	unsigned          tablet_slots_taken;
	unsigned          tablet_slot_counter;
	virg_tablet_meta *tablet_slots[VIRG_MEM_TABLETS];
	void             *gpu_slots;
	pthread_mutex_t   slot_lock;
	int               dbfd;
	unsigned          threads_per_block;
	unsigned          multi_threads;
	int               use_multi;
	int               use_gpu;
	int               use_stream;
	int               use_mmap;
} virginian;

/* reference: src/virginian.h:513-524 */
typedef struct {
	virginian        *v;
	virg_vm          *vm;
	virg_tablet_meta *tab;
	virg_tablet_meta *res;
	unsigned          row;
	unsigned          num_rows;
	unsigned          num_tablets;
	unsigned          tablets_proced;
	pthread_mutex_t   tab_lock;
	pthread_mutex_t   res_lock;
} virg_vm_arg;

/* reference: src/virginian.h:532-541 */
typedef struct {
	virg_vm          *vm;
	virg_tablet_meta *res;
	unsigned          row;
	char              buffer[VIRG_ROW_BUFFER];
} virg_reader;

static const size_t virg_testsizes[15] = {
	sizeof(int), sizeof(float), sizeof(long long int), sizeof(double),
	sizeof(char), sizeof(virg_tablet_meta), sizeof(virg_tablet_info),
	sizeof(virg_var), sizeof(virg_vm), sizeof(virg_op),
	sizeof(virg_result_node), sizeof(virg_db), sizeof(virg_reader),
	sizeof(virg_vm_arg), sizeof(virginian)
};

/* ---- slot manager / tablets / tables ---------------------------------- */
int virg_db_alloc(virginian *v, virg_tablet_meta **meta, int id);
/* Database files in the reference's format (reference: src/virginian.h:572-576,
 * src/db/create.c, open.c, close.c, clear.c, write.c).  virg_init leaves an
 * empty database in memory; virg_db_create(v, file) starts one backed by a
 * new file, virg_db_open(v, file) opens an existing one (before any tablet
 * is in memory), virg_db_close writes every tablet in memory and the meta
 * information.  Tablets must fit a disk slot (VIRG_TABLET_SIZE). */
int virg_db_create(virginian *v, const char *file);
int virg_db_open(virginian *v, const char *file);
int virg_db_close(virginian *v);
int virg_db_clear(virginian *v, unsigned slot);
int virg_db_write(virginian *v, unsigned slot);
int virg_db_load(virginian *v, unsigned tablet_id, virg_tablet_meta **tab);
int virg_db_loadnext(virginian *v, virg_tablet_meta **tab);
int virg_db_findslot(virginian *v, unsigned *slot);

int virg_table_create(virginian *v, const char *name, virg_t key_type);
int virg_table_addcolumn(virginian *v, unsigned table_id, const char *name, virg_t type);
int virg_table_insert(virginian *v, unsigned table_id, char *key, char *data, char *blob);
int virg_table_loadmem(virginian *v, unsigned table_id);
int virg_table_getid(virginian *v, const char *name, unsigned *id);
int virg_table_getcolumn(virginian *v, unsigned tid, const char *name, unsigned *id);
int virg_table_getcolumntype(virginian *v, unsigned tid, unsigned cid, virg_t *type);
int virg_table_getkeytype(virginian *v, unsigned tid, virg_t *type);
int virg_table_numrows(virginian *v, unsigned id, unsigned *rows);

int virg_tablet_addcolumn(virg_tablet_meta *tab, const char *name, virg_t type);
int virg_tablet_addrows(virginian *v, virg_tablet_meta *tab, unsigned rows);
int virg_tablet_addmaxrows(virginian *v, virg_tablet_meta *tab);
int virg_tablet_create(virginian *v, int *id, virg_t key_type, unsigned table_id);
int virg_tablet_check(virg_tablet_meta *t);
int virg_tablet_growfixed(virg_tablet_meta *tab, size_t size);
int virg_tablet_lock(virginian *v, unsigned tablet_id);
int virg_tablet_unlock(virginian *v, unsigned tablet_id);
int virg_tablet_addtail(virginian *v, virg_tablet_meta *head,
	virg_tablet_meta **tail, unsigned possible_rows);
int virg_tablet_remove(virginian *v, unsigned id);

/* ---- VM ---------------------------------------------------------------- */
int virg_vm_addop(virg_vm *vm, int op, int p1, int p2, int p3, virg_var p4);
int virg_vm_allocresult(virginian *v, virg_vm *vm, virg_tablet_meta **meta,
	virg_tablet_meta *template_);
int virg_vm_execute(virginian *v, virg_vm *vm);
int virg_vm_freeresults(virginian *v, virg_vm *vm);
virg_vm *virg_vm_init(void);
void virg_vm_cleanup(virginian *v, virg_vm *vm);

/* The drop-in boundary (reference: src/virginian.h:620-622,
 * src/vm/vm_gpu.cu:27-30,107).  Runs the data-parallel section of vm
 * (ops vm->pc .. Converge) over the tablet chain starting at tab[0] on the
 * GPU and fills host result tablets starting at res[0]. */
int virg_vm_gpu(virginian *v, virg_vm *vm, virg_tablet_meta **tab,
	virg_tablet_meta **res, unsigned num_tablets);
const size_t *virg_gpu_getsizes(void);

/* ---- results ----------------------------------------------------------- */
int virg_reader_init(virginian *v, virg_reader *r, virg_vm *vm);
int virg_reader_free(virginian *v, virg_reader *r);
int virg_reader_row(virginian *v, virg_reader *r);
int virg_reader_nexttablet(virginian *v, virg_reader *r);
int virg_reader_getrows(virginian *v, virg_reader *r, unsigned *rows);

/* ---- SQL and top level ------------------------------------------------- */
int virg_sql(virginian *v, const char *querystr, virg_vm *vm);
int virg_init(virginian *v);
int virg_close(virginian *v);
int virg_query(virginian *v, virg_reader **reader, const char *query);
void virg_release(virginian *v, virg_reader *reader);

size_t virg_sizeof(virg_t type);
const char *virg_opstring(int op);

extern struct timeval virg_starttime;
extern struct timeval virg_endtime;
void virg_timer_start(void);
double virg_timer_stop(void);

#ifdef __cplusplus
}
#endif
#endif
