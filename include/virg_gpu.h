/*
 * Entry points of the gfx950 executor that sit beside the drop-in boundary.
 *
 * The boundary itself is declared in virginian.h exactly as the reference
 * declares it:
 *   int virg_vm_gpu(virginian *v, virg_vm *vm, virg_tablet_meta **tab,
 *                   virg_tablet_meta **res, unsigned num_tablets);
 *       replaces src/vm/vm_gpu.cu:107 (declared at src/virginian.h:620-621)
 *   const size_t *virg_gpu_getsizes(void);
 *       replaces src/vm/vm_gpu.cu:27-30 (src/virginian.h:622)
 *
 * The functions below have no counterpart in the reference.  They keep a
 * query's table resident in HBM and run its data-parallel section any
 * number of times, which is how the reference's "data already on the GPU"
 * timings (src/vm/vm_gpu.cu:152-199 with the tablet cached) are reproduced
 * without a host round trip per tablet.  All are plain C: pointers, sizes
 * and the reference's own structs; VIRG_SUCCESS (1) / VIRG_FAIL (0).
 * A vg_query is used by one host thread at a time.
 */
#ifndef VIRG_GPU_H
#define VIRG_GPU_H

#include <stddef.h>
#include "virginian.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vg_query vg_query;

/* Prepare a generated program (Table, ResultColumn..., Parallel, ...,
 * Result, Converge, Finish: what virg_sql or virg_vm_addop produce) for
 * repeated execution over the table its Table op names.  The table's
 * tablets are copied to HBM once.  Output capacity is `capacity` rows
 * (0 = the table's row count).  Returns NULL with a message in err when
 * the program cannot run (no device, unsupported op, operand types that
 * differ, too many registers).  flags (0 or an OR of):
 *   VG_QUERY_LIGHT      for serving: the table's tablet descriptors are
 *                       shared with other queries, the query uses the
 *                       device's stream, its control block comes from a
 *                       pool and the two-pass arrays are made on first use;
 *   VG_QUERY_NO_OUTPUT  no output memory; give it with vg_query_set_outputs;
 *   VG_QUERY_INTERP     run on the two-pass interpreter even when the
 *                       program could have a specialised kernel.
 * Unless VG_QUERY_INTERP or the policy (vg_jit_policy) says otherwise, the
 * program's data-parallel section is compiled at prepare time into a
 * single-pass kernel of its own (hiprtc, cached by program structure, so
 * programs that differ only in constants share it): every column is read
 * from HBM once and the result is written in row order in one launch.  A
 * program it cannot take runs on the interpreter; the results are the
 * same bit for bit. */
#define VG_QUERY_LIGHT     1
#define VG_QUERY_NO_OUTPUT 2
#define VG_QUERY_INTERP    4
vg_query *vg_query_create(virginian *v, const virg_op *ops, unsigned num_ops,
	unsigned long long capacity, int flags, char *err, size_t errlen);
/* Results go to caller device memory from now on: out[j] has room for
 * `capacity` rows of result column j.  Refused (VIRG_FAIL) while a batch
 * (vg_batch, vg_pbatch) holds the query: a batch keeps a copy of each
 * query's launch arguments, outputs included. */
int vg_query_set_outputs(vg_query *q, void *const *out, unsigned long long capacity);
/* Bytes per row of each result column; returns the number of columns. */
unsigned vg_query_out_widths(vg_query *q, unsigned *widths);
/* Device memory for caller-managed outputs (a serving arena). */
void *vg_device_alloc(virginian *v, size_t bytes);
void vg_device_free(virginian *v, void *p);
/* Refused (with a message, the query left as it was) while a batch holds
 * the query: free the batch, or reload it without the query, first. */
void vg_query_free(vg_query *q);

/* Set when prepared queries get a specialised kernel (policy >= 0) and
 * return the previous policy: 0 never, 1 (the default, or the value of the
 * environment variable VG_JIT) queries that are neither light nor flagged
 * VG_QUERY_INTERP and whose tables hold at least 8192 rows per compute unit
 * (smaller ones cannot fill the device with the single pass's chunks), 2
 * every query not flagged VG_QUERY_INTERP. */
int vg_jit_policy(int policy);
/* What runs q: *fused = 1 for its specialised single-pass kernel (rows per
 * lane, tiles per chunk, workgroups, compile milliseconds; *mode bit 0: the
 * variant for 16-byte aligned columns, bit 1: filter columns staged in
 * LDS, bit 2: a one-column filter held in registers for both phases, bit
 * 3: pipelined (a workgroup counts its next chunk before it looks back for
 * the current one), bits 8-15: the waves per SIMD the kernel was compiled
 * for, 0 for the compiler's choice), 0 for the interpreter's count and
 * write passes. */
int vg_query_kernel(vg_query *q, int *fused, int *rows_per_lane, unsigned *tiles_per_chunk,
	unsigned *grid, double *compile_ms, int *mode);

/* Launch the section once on the query's stream (asynchronous). */
int vg_query_launch(vg_query *q);
/* Wait for the last launch; *rows = selected rows.  Fails when the result
 * exceeded the capacity. */
int vg_query_sync(vg_query *q, unsigned long long *rows);
/* launch + sync */
int vg_query_run(vg_query *q, unsigned long long *rows);
/* Copy the first min(last rows, max_rows) result rows into host arrays,
 * one per result column, in the reference's single-core row order. */
int vg_query_fetch(vg_query *q, void **cols, unsigned long long max_rows);
/* Mean milliseconds per launch over `iters` back-to-back launches after
 * `warmup` untimed ones, measured with HIP events on the query's stream. */
int vg_query_time(vg_query *q, int warmup, int iters, double *ms_per_launch,
	unsigned long long *rows);
/* Mean milliseconds of the count pass and of the write pass over `iters`
 * launches after `warmup` untimed ones, each launch bracketed by HIP
 * events on the query's stream and waited for before the next.  A query
 * on its specialised kernel has one pass: *ms_count is 0 and *ms_write the
 * whole launch. */
int vg_query_time_passes(vg_query *q, int warmup, int iters, double *ms_count,
	double *ms_write, unsigned long long *rows);
/* Rows scanned, tiles, referenced-column bytes per row, result bytes per
 * selected row, result columns. */
int vg_query_info(vg_query *q, unsigned long long *rows, unsigned *tiles,
	unsigned *read_bytes_per_row, unsigned *write_bytes_per_row, unsigned *nout);

/* A batch of prepared queries (distinct, on one device) replayed as one
 * HIP graph: each query's launch (its specialised kernel, or the count and
 * write passes) on its own forked stream, so small queries run side by
 * side and the host submits one graph per batch.  Results stay with each
 * query (vg_query_fetch after vg_batch_sync).  The queries must not be
 * launched by other means while the batch is in flight, and the batch
 * holds them (vg_query_free and vg_query_set_outputs refuse them) until
 * it is freed. */
typedef struct vg_batch vg_batch;
vg_batch *vg_batch_create(vg_query *const *qs, unsigned n, char *err, size_t errlen);
void vg_batch_free(vg_batch *b);
int vg_batch_launch(vg_batch *b);
/* Wait for the batch; rows[i] (when rows is not NULL) = rows query i selected. */
int vg_batch_sync(vg_batch *b, unsigned long long *rows);
/* Mean milliseconds per replay after `warmup` untimed ones (HIP events). */
int vg_batch_time(vg_batch *b, int warmup, int iters, double *ms);

/* A batch of prepared queries (distinct, on one device) run by persistent
 * launches over device-resident lists of (query, tile range) items: a
 * count launch whose workgroups take the items statically (workgroup g
 * items g, g + grid, ...), a scan launch turning each query's item counts
 * into output offsets, a write launch dealing out the write items the same
 * way, and a finish launch recording each query's outcome.  The single
 * pass (config.single) is one launch instead: items are chunks of a query's
 * tiles, taken in order through that query's own counter; a workgroup
 * counts its chunk, publishes the count, finds the chunk's output offset
 * by look-back over the query's earlier chunks and writes the chunk's rows
 * there, and the workgroup finishing a query's last chunk records its
 * outcome.  The batch holds its queries (vg_query_free and
 * vg_query_set_outputs refuse them) until it is freed or reloaded.  Queries are grouped by the occupancy their
 * register files allow, each group with its own pair of launches on its
 * own branch; the reset launch and every group's pair are captured as one
 * HIP graph, so a run is one graph launch whatever the number of queries.
 * Each query keeps its own results, exactly as a vg_query_run would leave
 * them (vg_query_fetch after vg_pbatch_sync).  The queries must not be
 * launched by other means while a run is in flight. */
typedef struct {
	unsigned tiles_per_item;   /* tiles (1024 rows) per work item; 0: 8 */
	unsigned grid_div;         /* launch 1/grid_div of the workgroups that fit; 0: 1 */
	unsigned split;            /* 0: group the queries by occupancy, 1: one group */
	unsigned order;            /* write list: 0 batch order, 1 smallest query first,
	                            * 2 largest first (count lists are largest first) */
	unsigned direct;           /* 1: launch directly instead of replaying a captured graph
	                            * (for batches whose queries change: vg_pbatch_reload) */
	virginian *db;             /* direct batches created empty: the database */
	unsigned single;           /* 1: one pass into each query's capacity as given (an
	                            * estimate): every item (tiles_per_item tiles) counts,
	                            * takes its output offset by decoupled look-back over
	                            * the query's earlier items and writes its rows there;
	                            * a query whose result does not fit reports it
	                            * (vg_pbatch_results) */
	unsigned jit;              /* 1: one persistent launch of specialised code: the
	                            * batch's programs compiled into one module at build
	                            * time (hiprtc), workgroups taking (query, chunk) items
	                            * in order from a device counter, each chunk read once
	                            * (filter columns staged in LDS), its output offset by
	                            * look-back over the query's earlier chunks; items,
	                            * order, grid and grouping settings do not apply */
} vg_pbatch_config;
typedef struct vg_pbatch vg_pbatch;
vg_pbatch *vg_pbatch_create(vg_query *const *qs, unsigned n, const vg_pbatch_config *cfg,
	char *err, size_t errlen);
void vg_pbatch_free(vg_pbatch *b);
/* Replace the batch's queries (after its last run). */
int vg_pbatch_reload(vg_pbatch *b, vg_query *const *qs, unsigned n, char *err, size_t errlen);
/* 1 when the batch's last run has finished, else 0 (does not wait). */
int vg_pbatch_done(vg_pbatch *b);
/* Rebuild the item lists and the graph for another configuration. */
int vg_pbatch_configure(vg_pbatch *b, const vg_pbatch_config *cfg, char *err, size_t errlen);
/* The configuration in use and the number of query groups. */
int vg_pbatch_get_config(vg_pbatch *b, vg_pbatch_config *cfg, unsigned *ngroups);
int vg_pbatch_launch(vg_pbatch *b);
/* Wait for the run.  rows[i]: rows query i selected; latency_ms[i]: when
 * its last write item finished, from the start of the run, by the device
 * clock.  Either may be NULL.  Fails when a result exceeded its capacity. */
int vg_pbatch_sync(vg_pbatch *b, unsigned long long *rows, double *latency_ms);
/* Wait for the run without failing on a query that did not fit: rows[i]
 * (rows query i selected, or at least its capacity), latency_ms[i] as
 * above, flags[i] = 0 when the result is complete, else 1 (capacity
 * exceeded) and/or 2 (a single-pass wait gave up).  Any may be NULL. */
int vg_pbatch_results(vg_pbatch *b, unsigned long long *rows, double *latency_ms, unsigned *flags);
/* Mean milliseconds per run over `iters` graph replays after `warmup`. */
int vg_pbatch_time(vg_pbatch *b, int warmup, int iters, double *ms);
/* Autotuner: times every candidate configuration (tiles per item 1..64,
 * the three write orders, full and half grids, grouped or not) on this
 * batch over `iters` replays each and keeps the fastest. */
int vg_pbatch_tune(vg_pbatch *b, int iters, vg_pbatch_config *best, double *best_ms,
	char *err, size_t errlen);
/* The plain way, for comparison: each query's count and write launches,
 * one query after another on one stream.  Mean milliseconds per round of
 * all n queries over `iters` rounds after `warmup`; latency_ms[i]
 * (optional): query i's results complete, from the start of the last
 * round (HIP events).  vg_query_sync then gives each query's rows. */
int vg_plain_time(vg_query *const *qs, unsigned n, int warmup, int iters, double *ms,
	double *latency_ms);

/* Out-of-core select: the table stays in host memory; tablets are taken in
 * batches of about batch_rows rows (0: 8M) and only the columns the
 * program loads are copied to HBM, on two streams in turn, overlapping
 * copies with kernels; result rows are appended to out[j] (one host array
 * per result column, room for `capacity` rows) in single-core row order.
 * The same path serves virg_vm_gpu when virginian.use_stream is set
 * (reference: src/vm/vm_gpu.cu:249-440). */
typedef struct {
	unsigned long long rows_scanned;
	unsigned long long rows_out;
	unsigned long long h2d_bytes;     /* column bytes copied to the device */
	unsigned long long d2h_bytes;     /* result bytes copied back */
	double seconds;                   /* wall time of the whole scan */
	unsigned batches;
} vg_stream_stats;
int vg_stream_select(virginian *v, const virg_op *ops, unsigned num_ops, void **out,
	unsigned long long capacity, unsigned long long batch_rows, vg_stream_stats *st,
	char *err, size_t errlen);

/* Tablet slot memory allocated since virg_init: pinned (hipHostMalloc)
 * and pageable (no device, or the system refused to pin more). */
int vg_host_memory(virginian *v, unsigned long long *pinned, unsigned long long *pageable);
/* Pin at most `bytes` of tablet slot memory from now on (0: no limit);
 * slots beyond it are pageable and reach the device through host-thread
 * copies into a pinned staging buffer. */
int vg_host_pin_limit(virginian *v, unsigned long long bytes);

/* Slot pager.  With a database file open, at most `tablets` tablets stay
 * in host memory (default VIRG_MEM_TABLETS = 64, as the reference's slot
 * table; 0 = no limit): loading or creating one more writes an unlocked
 * tablet to the file, chosen round-robin, and reuses its slot; a tablet not
 * in memory is read back from the file when it is next loaded (reference:
 * src/db/findslot.c:36-57, src/db/load.c:44-93).  Tablets larger than a
 * disk slot (VIRG_TABLET_SIZE) are never written out.  Pointers from
 * virg_table_tablets stay valid only while their tablets stay in memory;
 * the device paths copy what they need while a tablet is locked.
 * Returns the previous limit. */
unsigned vg_tablet_limit(virginian *v, unsigned tablets);
/* Tablets in host memory, tablets written out to make room and tablets
 * read back from the file since virg_init; *resident_id (if not NULL) = 1
 * when tablet `id` is in host memory. */
int vg_pager_stats(virginian *v, unsigned *resident, unsigned long long *evictions,
	unsigned long long *file_loads, unsigned id, int *resident_id);

/* Pinned host memory (results copy into it at link speed). */
void *vg_host_alloc(size_t bytes);
void vg_host_free(void *p);

/* Best host-to-device and device-to-host copy rates (GB/s) of one pinned
 * buffer of `bytes` over `reps` copies each. */
int vg_link_probe(virginian *v, size_t bytes, int reps, double *h2d_gbs, double *d2h_gbs);

/* Features of a (program, tablet) pair for the result-size predictor (no
 * counterpart in the reference), one row of VG_FEATURES floats per tablet:
 *   [0..2] logit(e_u) / VG_FEAT_SKIP for u = 1/2, 1/4, 3/4, where e_u is the
 *          fraction of the tablet's rows the filter keeps if its compares
 *          were independent, `column op constant` compares passing at the
 *          column histogram's estimate and any other compare at u
 *          (logit clipped to p in [1e-4, 1 - 1e-4]);
 *   [3]    e_{1/2};
 *   [4..10] compares estimated by histograms / 4, not estimated / 4, keep
 *          branches / 4, drop branches / 4, Invalid ops, logic ops,
 *          program length / 40;
 *   [11]   tablet rows / 2^20;
 *   [12..15] logit(e) / VG_FEAT_SKIP with the compares the histograms
 *          cannot estimate passing by kind: ordered (<, <=, >, >=) at 1/2,
 *          = at 0.05, != at 0.95; ordered at 1/4, the others at 1/2;
 *          ordered at 3/4, the others at 1/2; ordered at 1/2, = at 0,
 *          != at 1;
 *   [16..] the tablet's statistics as given (tabfeat).
 * Statistics per table: edges[c][0..BINS] of each fixed column's range,
 * per tablet hist[t][c][bin] (fractions), is_int[c], and tabfeat[t][...]
 * (VG_FEAT_COLS * (VG_FEAT_BINS + 2) values). */
#define VG_FEAT_COLS 6
#define VG_FEAT_BINS 16
#define VG_FEAT_SKIP 4.0
#define VG_FEATURES (16 + VG_FEAT_COLS * (VG_FEAT_BINS + 2))
typedef struct {
	unsigned ntab;
	const double *rows;
	const double *edges;
	const double *hist;
	const int *is_int;
	const double *tabfeat;
} vg_table_stats;
int vg_program_features(const virg_op *ops, unsigned num_ops, const vg_table_stats *st, float *out);

/* Result-size predictor (no counterpart in the reference): a 3-layer
 * perceptron nin -> nh -> nh -> 1 (ReLU, ReLU, sigmoid; f32) predicting
 * the fraction of a tablet's rows a program selects, trained on the GPU
 * by its own forward, backward and Adam kernels.  Parameters, flat f32:
 * W1[nin][nh] b1[nh] W2[nh][nh] b2[nh] W3[nh] b3; loss = mean (y - t)^2.
 * With skip != 0 the output pre-activation also gets skip * x[0] (no
 * parameter) and W3 starts at 0, so the network corrects an estimate
 * whose logit / skip is fed as feature 0.
 * Arrays are host memory, examples row-major (n x nin). */
typedef struct vg_mlp vg_mlp;
vg_mlp *vg_mlp_create(unsigned nin, unsigned nh, float skip, unsigned long long seed);
void vg_mlp_free(vg_mlp *m);
unsigned vg_mlp_num_params(const vg_mlp *m);
int vg_mlp_get_params(const vg_mlp *m, float *out);
int vg_mlp_set_params(vg_mlp *m, const float *in);
int vg_mlp_forward(vg_mlp *m, const float *x, unsigned n, float *y);
/* The same on the host, from a copy of the parameters taken whenever they
 * change: no device call, for callers that must not wait for the device. */
int vg_mlp_forward_host(const vg_mlp *m, const float *x, unsigned n, float *y);
int vg_mlp_gradient(vg_mlp *m, const float *x, const float *t, unsigned n, float *loss, float *grad);
/* The loss vg_mlp_gradient and vg_mlp_train use: 0 mean squared error
 * (the default), 1 mean absolute error. */
int vg_mlp_set_loss(vg_mlp *m, int kind);
/* The pinball loss of quantile tau, mean over examples of
 * max(tau (t - y), (1 - tau) (y - t)) (0 < tau < 1): the network learns
 * the tau-quantile of the target, an upper bound where tau is high. */
int vg_mlp_set_quantile(vg_mlp *m, float tau);
int vg_mlp_train(vg_mlp *m, const float *x, const float *t, unsigned n, unsigned epochs, unsigned batch,
	float lr, unsigned long long seed, float *loss);

/* Rows the last launch of q selected in each of its tablets (n entries at
 * most, in table order; the per-tablet counts the predictor learns from);
 * returns the number of tablets. */
unsigned vg_query_tablet_rows(vg_query *q, unsigned long long *rows, unsigned n);

/* Architecture name and compute-unit count of the device in use
 * (VIRG_FAIL and *cus = 0 without one). */
int vg_device_info(virginian *v, char *name, size_t n, int *cus);

/* Bulk table builder on the device: table `name` with the standard
 * synthetic schema (int64 key = row index; uniformi, normali5, normali20
 * int32; uniformf, normalf5, normalf20 f32) and `rows` rows in host tablets
 * of `tablet_rows` rows, identical bit for bit to what the host generator
 * builds for the same arguments.  Replaces, for bulk loads, the reference's
 * row-at-a-time ingest (virg_table_insert, reference src/table/insert.c;
 * src/tablet/addrows.c:39-73).  Needs a device; tablet_rows > 0. */
int vg_gen_table(virginian *v, const char *name, unsigned long long rows, unsigned tablet_rows,
	unsigned seed);

/* Decode the data-parallel section starting at op `pc` against tablet
 * `tab`'s schema into the kernel's micro-program (no device needed).
 * `out` receives a vg_prog (see virginian_amd/csrc/gpu/vg_prog.h). */
int vg_decode_program(const virg_op *ops, unsigned num_ops, unsigned pc,
	const virg_tablet_meta *tab, void *out, size_t out_bytes, char *err, size_t errlen);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops and harnesses that
 * forbid hipMalloc).  Runs the data-parallel section of a generated program
 * (ops as virg_sql or virg_vm_addop make them) over `rows` rows already in
 * device memory -- cols[c] is column c's row 0, col_types[c] its virg_t,
 * key the key column (read only by Rowid; may be NULL otherwise) -- into
 * the caller's device outputs out[j] (room for `capacity` rows of result
 * column j), on `stream`, asynchronously, with the interpreter's count and
 * write passes (the results equal vg_query_run's bit for bit).  result is a
 * 32-byte device block: bytes 0-7 the rows selected, bytes 8-11 nonzero
 * when they exceeded capacity.  The call allocates nothing: its scratch is
 * a fixed array of about 4.4 MB per GPU that the library holds until the
 * process ends, so rows beyond 16,777,216 are refused.  Calls that overlap
 * on one GPU share that array and race.  HIP errors on the stream are the
 * caller's to check.  Declared here with no stream type: stream is a
 * hipStream_t. */
int vg_select_noalloc(const virg_op *ops, unsigned num_ops, const void *const *cols,
	const int *col_types, unsigned ncols, const void *key, int key_type, unsigned long long rows,
	void *const *out, unsigned long long capacity, void *result, void *stream, char *err, size_t errlen);

/* Serving loop (no counterpart in the reference, whose queries run one at
 * a time: src/vm/execute.c:28-102).  Requests are generated op-programs
 * over tables already loaded; they arrive over time and are served with
 * continuous batching by one host thread:
 *   - a request joins the waiting line once the server's clock passes its
 *     arrival time;
 *   - while a runner (a stream) is free, waiting requests are admitted in
 *     arrival order (reruns first) as long as their result buffers fit in
 *     the arena (device memory taken once at creation): admission by HBM
 *     footprint;
 *   - while the line is short (below batch_line) each run takes one
 *     request, run alone as its count and write launches (a last launch
 *     stores the outcome into pinned host memory, which the loop polls),
 *     and at most one costly request (rows x ops, see `costly`) runs at a
 *     time: costly runs side by side only slow each other, while cheap ones
 *     beside a costly one barely do; from batch_line on, runs take up to
 *     max_batch requests as one persistent batch (vg_pbatch) whose kernels
 *     store every outcome into pinned memory; runs on different runners
 *     overlap on the device;
 *   - a finished run completes its requests, except a result that did not
 *     fit its buffers: that request is admitted again, before new ones, with
 *     buffers of the exact size the run counted.
 * Result buffers are sized (`sizing`) for every row of the query's table,
 * or per tablet as rows * min(1, max(p, 0) + margin), f32, summed, plus
 * 64, where p is the independence estimate of the filter over the table's
 * histograms (feature 3 of vg_program_features) or the result-size
 * network's output.  Results never depend on the sizing.  With runners = 1
 * and max_batch = 1 this is the plain way: each request alone, in arrival
 * order, waited for, then the next.  Latency of a request = when the loop
 * saw its result complete in HBM (`ready`) minus its arrival.  A server is
 * used by one host thread at a time. */
#define VG_SIZE_ALL      0   /* buffers for every row of the table */
#define VG_SIZE_ESTIMATE 1   /* the independence estimate over the histograms */
#define VG_SIZE_NET      2   /* the result-size network (vg_mlp) */
typedef struct {
	unsigned long long arena_bytes;   /* result arena (0: 8 GiB) */
	unsigned runners;                 /* runs in flight at once, one stream each (0: 3) */
	unsigned max_batch;               /* requests per run (0: 64) */
	unsigned tiles_per_item;          /* batch work items (0: 8) */
	unsigned sizing;                  /* VG_SIZE_* */
	float margin;                     /* added to each tablet's predicted fraction */
	unsigned single;                  /* 1: single-pass batches (every run a batch) */
	unsigned split;                   /* vg_pbatch_config.split */
	unsigned grid_div;                /* vg_pbatch_config.grid_div */
	unsigned keep;                    /* 1: each result is copied to host memory the server
	                                   * holds (until vg_server_release) before the request
	                                   * counts as done */
	const vg_mlp *net;                /* VG_SIZE_NET: the trained network */
	unsigned batch_line;              /* waiting requests from which runs are batches; below
	                                   * it each run takes one request (0: 16) */
	unsigned long long costly;        /* rows x ops from which a request is costly: while the
	                                   * line is short at most one costly run is in flight,
	                                   * cheaper requests may pass a waiting costly one
	                                   * (0: 1e8; ~0: no limit) */
	unsigned host_arena;              /* 1: the arena is pinned host memory the kernels write
	                                   * results into directly (no copy to the host; the
	                                   * writes cross the link) */
} vg_server_config;
typedef struct {
	unsigned long long done, failed, reruns;
	unsigned long long lone;          /* runs of one request */
	unsigned long long batches;       /* runs of several */
	unsigned long long reserved_bytes;/* result bytes reserved, over all admissions */
	unsigned long long arena_peak;    /* most arena bytes in use at once */
	unsigned long long in_flight_peak;
	double admit_s, launch_s, collect_s;  /* host seconds per step of the loop */
} vg_server_stats;
/* state: 0 not arrived, 1 waiting, 2 running, 3 done, 4 failed */
typedef struct {
	unsigned state;
	unsigned reruns;
	unsigned long long rows;
	unsigned long long capacity;      /* result rows reserved for its last run */
	double arrival, ready;            /* seconds on the server's clock */
} vg_server_req;
typedef struct vg_server vg_server;
vg_server *vg_server_create(virginian *v, const vg_server_config *cfg, char *err, size_t errlen);
void vg_server_free(vg_server *s);
/* Statistics of a table's tablets for sizing by estimate or network
 * (copied). */
int vg_server_table(vg_server *s, unsigned table_id, const vg_table_stats *st);
/* Queue a request arriving at `arrival` seconds on the server's clock
 * (negative: now); *id numbers requests from 0 in submission order.  The
 * ops are copied (ResultColumn names are not kept). */
int vg_server_submit(vg_server *s, const virg_op *ops, unsigned num_ops, double arrival,
	unsigned long long *id);
/* One pass of the loop (never waits); the number of requests completed,
 * or -1 when a launch failed. */
int vg_server_step(vg_server *s);
/* Restart the clock at 0 and serve until every submitted request is done
 * or failed (timeout_s > 0: give up after that long). */
int vg_server_run(vg_server *s, double timeout_s);
double vg_server_now(vg_server *s);
int vg_server_request(vg_server *s, unsigned long long id, vg_server_req *out);
/* keep = 1: a done request's result rows into host arrays, one per result
 * column, in the reference's single-core row order (as vg_query_fetch);
 * release drops the server's copy. */
int vg_server_fetch(vg_server *s, unsigned long long id, void **cols, unsigned long long max_rows);
int vg_server_release(vg_server *s, unsigned long long id);
int vg_server_get_stats(vg_server *s, vg_server_stats *out);
/* The last error's message ("" when none). */
const char *vg_server_error(vg_server *s);

#ifdef __cplusplus
}
#endif

#endif
