/*
 * Device layer: HIP context, pinned host slots, the device-resident tablet
 * cache, the virg_vm_gpu boundary and the vg_query C ABI.
 *
 * Differences from the reference host driver (src/vm/vm_gpu.cu:107-578):
 *  - one pair of launches (count pass, write pass) covers every tablet of
 *    the query instead of one launch per tablet with a full 8 MB result
 *    copy after each;
 *  - tablets are copied to HBM once and stay resident (keyed by tablet id,
 *    re-copied when their row count or layout changes), so repeated
 *    queries read them from HBM; the reference's serial / stream / mapped
 *    transfer modes are all served by this cache;
 *  - the program is decoded on the host and passed as a kernel argument
 *    instead of copying virg_vm to __constant__ memory;
 *  - results are produced in row order and copied back only as large as
 *    they are, filling each host result tablet to capacity.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <stdarg.h>
#include <algorithm>
#include <atomic>
#include <mutex>
#include <chrono>
#include <thread>
#include <unordered_map>
#include <vector>

extern "C" {
#include "../host/internal.h"
#include "vg_prog.h"
#include "vgx.h"
#include "virg_gpu.h"
}

extern "C" int vg_kernel_launch(const vg_prog *P, vg_launch *L, hipStream_t stream, hipEvent_t mid);
extern "C" uint32_t vg_kernel_tile_rows(void);
extern "C" uint32_t vg_kernel_grid(const vg_prog *P, uint32_t num_tiles);
extern "C" int vg_kernel_check(const vg_prog *P);
extern "C" int vg_decode(const virg_op *stmt, unsigned num_ops, unsigned pc,
	const virg_tablet_meta *tab, const virg_tablet_meta *res, vg_prog *P, char *err, size_t errlen);

/* validity bytes per tile: one per lane of the 256-thread write workgroup */
#define VG_VBYTES_PER_TILE 256u

struct resident {
	void *dptr = nullptr;
	size_t bytes = 0;
	unsigned rows = 0;
	size_t size = 0, fixed_block = 0;
	unsigned possible_rows = 0;
};

/* launch resources owned by one query stream: control block, validity
 * bytes and per-wave counts of every tile, per-workgroup counts */
struct launch_res {
	vg_ctrl *ctrl = nullptr;
	uint8_t *vbits = nullptr;
	uint32_t *tcount = nullptr;
	size_t tiles_cap = 0;
	uint32_t *wgcount = nullptr;
	size_t wgcount_cap = 0;
};

/* device description of a list of resident tablets */
struct plan {
	std::vector<vg_tabdesc> desc;
	std::vector<uint32_t> tile_tab;
	vg_tabdesc *d_desc = nullptr;
	uint32_t *d_tile = nullptr;
	uint32_t num_tiles = 0;
	uint64_t rows = 0;
};

/* a plan shared by the queries over one table (keyed by its tablets) */
struct shared_plan {
	plan p;
	std::vector<std::pair<unsigned, unsigned>> key;    /* (tablet id, rows) */
};

struct vgd_device {
	int device = 0;
	hipStream_t stream = nullptr;
	std::unordered_map<unsigned, resident> tablets;
	launch_res lr;
	void *outbuf = nullptr;
	size_t outbuf_bytes = 0;
	hipDeviceProp_t prop;
	std::unordered_map<unsigned, shared_plan *> plans;  /* by table id */
	std::vector<shared_plan *> retired;                 /* freed at close */
	std::vector<vg_ctrl *> ctrl_free;                   /* pooled control blocks */
	std::vector<vg_ctrl *> ctrl_chunks;
	std::vector<launch_res> lr_free;                    /* pooled two-pass arrays */
	struct stream_cache *scache = nullptr;              /* streaming buffers, kept between calls */
};
static void free_stream_cache(vgd_device *d);

static void free_plan(plan &p);
static int build_plan(vgd_device *d, const std::vector<const virg_tablet_meta *> &tabs,
	hipStream_t stream, plan &p);

static int herr(hipError_t e, const char *what)
{
	if (e != hipSuccess) {
		fprintf(stderr, "::: Virginian HIP error in %s: %s\n", what, hipGetErrorString(e));
		return 1;
	}
	return 0;
}
#define HCHK(call, what) do { if (herr((call), (what))) return VIRG_FAIL; } while (0)

static vgd_device *dev_of(virginian *v)
{
	virg_state *s = virg_state_of(v);
	return s ? s->dev : nullptr;
}

/* ---------------------------------------------- allocation-free launches */
/* Scratch of vg_select_noalloc: the launch's tablet descriptor and tile
 * table, validity bytes and per-wave counts of every tile, per-workgroup
 * counts.  One array per GPU the library runs on, held until the process
 * ends.  Host code never uses the array's name as an address: the device
 * address of each GPU's copy comes from hipGetSymbolAddress, once per
 * device with that device current (noalloc_scratch). */
#define VG_NOALLOC_MAX_ROWS  (16u << 20)
#define VG_NOALLOC_MAX_TILES (VG_NOALLOC_MAX_ROWS / 1024u)
#define VG_NA_DESC    0u
#define VG_NA_TILES   256u
#define VG_NA_VBITS   (VG_NA_TILES + VG_NOALLOC_MAX_TILES * 4u)
#define VG_NA_TCOUNT  (VG_NA_VBITS + VG_NOALLOC_MAX_TILES * 256u)
#define VG_NA_WGCOUNT (VG_NA_TCOUNT + VG_NOALLOC_MAX_TILES * 16u)
#define VG_NA_BYTES   (VG_NA_WGCOUNT + VG_NOALLOC_MAX_TILES * 4u)
static __device__ __align__(256) uint8_t scratch[VG_NA_BYTES];

#define VG_NA_MAX_DEVICES 64
static std::mutex g_na_mu;
static uint8_t *g_na_ptr[VG_NA_MAX_DEVICES];

/* Device address of the current device's scratch (looked up once per
 * device); NULL if the lookup fails. */
static uint8_t *noalloc_scratch(int dev)
{
	if (dev < 0 || dev >= VG_NA_MAX_DEVICES)
		return nullptr;
	std::lock_guard<std::mutex> g(g_na_mu);
	if (!g_na_ptr[dev]) {
		void *p = nullptr;
		if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
			(void)hipGetLastError();
			return nullptr;
		}
		g_na_ptr[dev] = (uint8_t *)p;
	}
	return g_na_ptr[dev];
}

extern "C" int vgd_open(virginian *v)
{
	virg_state *s = virg_state_of(v);
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
		(void)hipGetLastError();
		s->dev = nullptr;
		return 0;
	}
	vgd_device *d = new vgd_device();
	d->device = VIRG_CUDADEVICE;
	if (hipSetDevice(d->device) != hipSuccess ||
		hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess ||
		hipGetDeviceProperties(&d->prop, d->device) != hipSuccess) {
		(void)hipGetLastError();
		delete d;
		s->dev = nullptr;
		return 0;
	}
	s->dev = d;
	(void)noalloc_scratch(d->device);     /* set up while allocation is allowed */
	return 1;
}

static void free_lr(launch_res &lr)
{
	if (lr.ctrl) (void)hipFree(lr.ctrl);
	if (lr.vbits) (void)hipFree(lr.vbits);
	if (lr.tcount) (void)hipFree(lr.tcount);
	if (lr.wgcount) (void)hipFree(lr.wgcount);
	lr = launch_res();
}

extern "C" void vgd_close(virginian *v)
{
	virg_state *s = virg_state_of(v);
	if (!s || !s->dev)
		return;
	vgd_device *d = s->dev;
	(void)hipStreamSynchronize(d->stream);
	for (auto &kv : d->tablets)
		(void)hipFree(kv.second.dptr);
	free_lr(d->lr);
	if (d->outbuf)
		(void)hipFree(d->outbuf);
	for (auto &kv : d->plans)
		d->retired.push_back(kv.second);
	for (shared_plan *sp : d->retired) {
		free_plan(sp->p);
		delete sp;
	}
	for (vg_ctrl *c : d->ctrl_chunks)
		(void)hipFree(c);
	for (launch_res &lr : d->lr_free)
		free_lr(lr);
	free_stream_cache(d);
	(void)hipStreamDestroy(d->stream);
	delete d;
	s->dev = nullptr;
}

extern "C" void vgd_forget_tablet(virginian *v, unsigned id)
{
	vgd_device *d = dev_of(v);
	if (!d)
		return;
	auto it = d->tablets.find(id);
	if (it == d->tablets.end())
		return;
	/* shared plans may hold the tablet's address: build them again */
	for (auto &kv : d->plans)
		d->retired.push_back(kv.second);
	d->plans.clear();
	(void)hipDeviceSynchronize();
	(void)hipFree(it->second.dptr);
	d->tablets.erase(it);
}

extern "C" void *vgd_host_alloc(size_t bytes)
{
	void *p = nullptr;
	if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
		(void)hipGetLastError();
		return nullptr;
	}
	return p;
}

/* Frees slot memory from vgd_host_alloc, or from the plain allocation the
 * host library falls back to when pinning fails. */
extern "C" void vgd_host_free(void *p)
{
	if (hipHostFree(p) != hipSuccess) {
		(void)hipGetLastError();
		free(p);
	}
}

/* Make a tablet resident in HBM; returns its device copy. */
static void *resident_tablet(vgd_device *d, const virg_tablet_meta *t, hipStream_t stream)
{
	auto it = d->tablets.find(t->id);
	if (it != d->tablets.end()) {
		resident &r = it->second;
		if (r.rows == t->rows && r.size == t->size && r.fixed_block == t->fixed_block &&
			r.possible_rows == t->possible_rows)
			return r.dptr;
		(void)hipDeviceSynchronize();
		(void)hipFree(r.dptr);
		d->tablets.erase(it);
	}
	resident r;
	r.bytes = t->size;
	if (hipMalloc(&r.dptr, r.bytes) != hipSuccess) {
		(void)hipGetLastError();
		fprintf(stderr, "::: Virginian: cannot make tablet %u (%zu bytes) resident in HBM\n", t->id, r.bytes);
		return nullptr;
	}
	if (hipMemcpyAsync(r.dptr, t, r.bytes, hipMemcpyHostToDevice, stream) != hipSuccess) {
		(void)hipGetLastError();
		(void)hipFree(r.dptr);
		return nullptr;
	}
	r.rows = t->rows;
	r.size = t->size;
	r.fixed_block = t->fixed_block;
	r.possible_rows = t->possible_rows;
	d->tablets[t->id] = r;
	return r.dptr;
}

/* Extra locks taken on tablets a walk has moved past, released when the
 * walk's user is done (the slot pager leaves locked tablets in place). */
struct held_locks {
	virginian *v;
	std::vector<unsigned> ids;
	explicit held_locks(virginian *v_) : v(v_) {}
	int take(unsigned id)
	{
		if (virg_tablet_lock(v, id) == VIRG_FAIL)
			return VIRG_FAIL;
		ids.push_back(id);
		return VIRG_SUCCESS;
	}
	~held_locks()
	{
		for (unsigned id : ids)
			virg_tablet_unlock(v, id);
	}
};

static int ensure_lr(launch_res &lr, uint32_t num_tiles, uint32_t grid, hipStream_t stream)
{
	if (!lr.ctrl) {
		HCHK(hipMalloc((void **)&lr.ctrl, sizeof(vg_ctrl)), "control block");
		HCHK(hipMemsetAsync(lr.ctrl, 0, sizeof(vg_ctrl), stream), "control block");
	}
	if (lr.tiles_cap < num_tiles) {
		if (lr.vbits) {
			HCHK(hipStreamSynchronize(stream), "validity bytes");
			HCHK(hipFree(lr.vbits), "validity bytes");
			HCHK(hipFree(lr.tcount), "tile counts");
			lr.vbits = nullptr;
			lr.tcount = nullptr;
		}
		size_t cap = num_tiles < 64 ? 64 : num_tiles;
		HCHK(hipMalloc((void **)&lr.vbits, cap * VG_VBYTES_PER_TILE), "validity bytes");
		HCHK(hipMalloc((void **)&lr.tcount, cap * 4 * sizeof(uint32_t)), "tile counts");
		lr.tiles_cap = cap;
	}
	if (lr.wgcount_cap < grid) {
		if (lr.wgcount) {
			HCHK(hipStreamSynchronize(stream), "counts");
			HCHK(hipFree(lr.wgcount), "counts");
		}
		size_t cap = grid < 4096 ? 4096 : grid;
		HCHK(hipMalloc((void **)&lr.wgcount, cap * sizeof(uint32_t)), "counts");
		lr.wgcount_cap = cap;
	}
	return VIRG_SUCCESS;
}

static void free_plan(plan &p)
{
	if (p.d_desc) (void)hipFree(p.d_desc);
	if (p.d_tile) (void)hipFree(p.d_tile);
	p.d_desc = nullptr;
	p.d_tile = nullptr;
}

static int build_plan(vgd_device *d, const std::vector<const virg_tablet_meta *> &tabs,
	hipStream_t stream, plan &p)
{
	const uint32_t T = vg_kernel_tile_rows();
	p.desc.clear();
	p.tile_tab.clear();
	p.rows = 0;
	for (size_t i = 0; i < tabs.size(); i++) {
		const virg_tablet_meta *t = tabs[i];
		const virg_tablet_meta *t0 = tabs[0];
		VIRG_CHECK(t->fixed_columns != t0->fixed_columns || t->key_type != t0->key_type ||
			memcmp(t->fixed_type, t0->fixed_type, sizeof t->fixed_type) != 0,
			"tablets of one table differ in schema");
		VIRG_CHECK(t->fixed_columns > VG_MAX_COLS - 1, "too many columns");
		char *base = (char *)resident_tablet(d, t, stream);
		if (!base)
			return VIRG_FAIL;
		vg_tabdesc td;
		memset(&td, 0, sizeof td);
		bool aligned = true;
		for (unsigned c = 0; c < t->fixed_columns; c++) {
			uint64_t a = (uint64_t)(base + t->fixed_block + t->fixed_offset[c]);
			td.col[c] = a;
			aligned = aligned && (a % 16 == 0);
		}
		uint64_t ka = (uint64_t)(base + t->key_block);
		td.col[VG_KEY_SRC] = ka;
		aligned = aligned && (ka % 16 == 0);
		td.rows = t->rows;
		td.first_tile = (uint32_t)p.tile_tab.size();
		td.aligned = aligned ? 1u : 0u;
		uint32_t nt = (t->rows + T - 1) / T;
		for (uint32_t k = 0; k < nt; k++)
			p.tile_tab.push_back((uint32_t)i);
		p.desc.push_back(td);
		p.rows += t->rows;
	}
	p.num_tiles = (uint32_t)p.tile_tab.size();
	free_plan(p);
	if (!p.desc.empty()) {
		HCHK(hipMalloc((void **)&p.d_desc, p.desc.size() * sizeof(vg_tabdesc)), "descriptors");
		HCHK(hipMemcpyAsync(p.d_desc, p.desc.data(), p.desc.size() * sizeof(vg_tabdesc),
			hipMemcpyHostToDevice, stream), "descriptors");
	}
	if (p.num_tiles) {
		HCHK(hipMalloc((void **)&p.d_tile, p.tile_tab.size() * sizeof(uint32_t)), "tile table");
		HCHK(hipMemcpyAsync(p.d_tile, p.tile_tab.data(), p.tile_tab.size() * sizeof(uint32_t),
			hipMemcpyHostToDevice, stream), "tile table");
	}
	return VIRG_SUCCESS;
}

static int launch(const vg_prog &P, const plan &p, launch_res &lr, void *const *out,
	uint64_t capacity, hipStream_t stream, hipEvent_t mid = nullptr, vg_bresult *hres = nullptr)
{
	if (p.num_tiles == 0) {
		if (mid)
			HCHK(hipEventRecord(mid, stream), "event");
		return VIRG_SUCCESS;
	}
	if (ensure_lr(lr, p.num_tiles, vg_kernel_grid(&P, p.num_tiles), stream) == VIRG_FAIL)
		return VIRG_FAIL;
	vg_launch L;
	memset(&L, 0, sizeof L);
	L.tabs = p.d_desc;
	L.tile_tab = p.d_tile;
	L.num_tiles = p.num_tiles;
	L.capacity = capacity;
	L.ctrl = lr.ctrl;
	L.vbits = lr.vbits;
	L.tcount = lr.tcount;
	L.wgcount = lr.wgcount;
	L.hres = hres;
	for (unsigned j = 0; j < P.nout; j++)
		L.out[j] = out[j];
	int rc = vg_kernel_launch(&P, &L, stream, mid);
	if (rc != 0) {
		fprintf(stderr, "::: Virginian: kernel launch failed (%d)\n", rc);
		return VIRG_FAIL;
	}
	return VIRG_SUCCESS;
}

static int read_ctrl(const launch_res &lr, hipStream_t stream, uint64_t *rows)
{
	vg_ctrl c;
	HCHK(hipMemcpyAsync(&c, lr.ctrl, sizeof c, hipMemcpyDeviceToHost, stream), "control block");
	HCHK(hipStreamSynchronize(stream), "kernel");
	if (c.overflow != 0) {
		/* leave a clean control block for the next launch */
		HCHK(hipMemsetAsync(lr.ctrl, 0, sizeof(vg_ctrl), stream), "control block");
		HCHK(hipStreamSynchronize(stream), "control block");
	}
	if (c.overflow != 0)
		fprintf(stderr, "::: Virginian: overflow flags 0x%x (1: output capacity, 2: a single-pass wait gave up)\n",
			c.overflow);
	VIRG_CHECK(c.overflow != 0, "result exceeded the output capacity");
	rows[0] = c.total;
	return VIRG_SUCCESS;
}

/* ---------------------------------------------------- out-of-core stream */
/* The table stays in (pinned) host memory.  Tablets are taken in batches;
 * for each batch only the columns the program loads are copied to HBM, on
 * two streams in turn, so the host-to-device copy of one batch overlaps the
 * kernels of the other and the device-to-host copy of earlier results (on
 * a third stream).  Results are appended in batch order, so the logical
 * result is the single-core row order.  The reference's streaming mode
 * (src/vm/vm_gpu.cu:249-440, v->use_stream) copies whole tablets and one
 * 8 MB result tablet per data tablet; here only referenced column bytes
 * cross the link and result copies are exactly as large as the results. */

struct sbuf {
	hipStream_t stream = nullptr;
	hipEvent_t done = nullptr;      /* kernels and control copy of the batch */
	hipEvent_t freed = nullptr;     /* the batch's results copied out */
	char *stage = nullptr;
	size_t stage_cap = 0;
	char *h_stage = nullptr;        /* pinned host copy of `stage` for pageable tablets */
	size_t h_stage_cap = 0;
	char *out = nullptr;
	size_t out_cap = 0;
	vg_tabdesc *d_desc = nullptr, *h_desc = nullptr;
	uint32_t *d_tile = nullptr, *h_tile = nullptr;
	size_t desc_cap = 0, tile_cap = 0;
	vg_ctrl *h_ctrl = nullptr;
	launch_res lr;
	bool busy = false;
	uint64_t rows = 0;
};

static void free_sbuf(sbuf &b)
{
	if (b.stream) (void)hipStreamSynchronize(b.stream);
	if (b.stage) (void)hipFree(b.stage);
	if (b.h_stage) (void)hipHostFree(b.h_stage);
	if (b.out) (void)hipFree(b.out);
	if (b.d_desc) (void)hipFree(b.d_desc);
	if (b.d_tile) (void)hipFree(b.d_tile);
	if (b.h_desc) (void)hipHostFree(b.h_desc);
	if (b.h_tile) (void)hipHostFree(b.h_tile);
	if (b.h_ctrl) (void)hipHostFree(b.h_ctrl);
	free_lr(b.lr);
	if (b.done) (void)hipEventDestroy(b.done);
	if (b.freed) (void)hipEventDestroy(b.freed);
	if (b.stream) (void)hipStreamDestroy(b.stream);
	b = sbuf();
}

/* A device's streaming buffers, kept from one streaming call to the next
 * (each grows to the largest batch seen and is held until the instance
 * closes): allocating and pinning them again on every call cost several
 * milliseconds a query. */
struct stream_cache {
	sbuf B[2];
	hipStream_t down = nullptr;     /* result copies */
};

static void free_stream_cache(vgd_device *d)
{
	if (!d->scache)
		return;
	for (int k = 0; k < 2; k++)
		free_sbuf(d->scache->B[k]);
	if (d->scache->down) {
		(void)hipStreamSynchronize(d->scache->down);
		(void)hipStreamDestroy(d->scache->down);
	}
	delete d->scache;
	d->scache = nullptr;
}

/* receives each batch's results, in order: n rows of every result column
 * at cols[j] (device memory, widths[j] bytes a row), copied on `down` */
struct stream_sink {
	virtual int take(void *const *cols, const size_t *widths, unsigned ncol, uint64_t n,
		hipStream_t down) = 0;
	virtual ~stream_sink() {}
};

static size_t align256(size_t x)
{
	return (x + 255) & ~(size_t)255;
}

/* host address of column c of a host tablet (VG_KEY_SRC: the key) */
static const char *host_column(const virg_tablet_meta *t, unsigned src)
{
	return src == VG_KEY_SRC ? (const char *)t + t->key_block
	                         : (const char *)t + t->fixed_block + t->fixed_offset[src];
}

/* true when host memory at p is pinned (page-locked, device-visible) */
static bool host_pinned(const void *p)
{
	hipPointerAttribute_t a;
	if (hipPointerGetAttributes(&a, p) != hipSuccess) {
		(void)hipGetLastError();
		return false;
	}
	return a.type == hipMemoryTypeHost;
}

/* host-to-host copies done by a pool of threads, in 4 MB pieces */
struct host_copy {
	char *dst;
	const char *src;
	size_t bytes;
};

static void parallel_copy(const std::vector<host_copy> &jobs)
{
	const size_t piece = 4u << 20;
	std::vector<host_copy> pieces;
	for (const host_copy &j : jobs)
		for (size_t o = 0; o < j.bytes; o += piece)
			pieces.push_back({j.dst + o, j.src + o, j.bytes - o < piece ? j.bytes - o : piece});
	unsigned nt = std::thread::hardware_concurrency();
	nt = nt < 1 ? 1 : nt > 16 ? 16 : nt;
	if (pieces.size() < nt)
		nt = (unsigned)pieces.size();
	std::atomic<size_t> next{0};
	auto work = [&]() {
		for (size_t i; (i = next.fetch_add(1)) < pieces.size();)
			memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes);
	};
	std::vector<std::thread> th;
	for (unsigned t = 1; t < nt; t++)
		th.emplace_back(work);
	work();
	for (std::thread &t : th)
		t.join();
}

static int stream_finish(vgd_device *d, const vg_prog &P, sbuf &b, stream_sink &sink,
	hipStream_t down, uint64_t *rows_out, uint64_t *d2h)
{
	HCHK(hipEventSynchronize(b.done), "stream batch");
	VIRG_CHECK(b.h_ctrl->overflow != 0, "stream batch overflowed its output");
	const uint64_t n = b.h_ctrl->total;
	void *cols[VG_MAX_OUT];
	size_t widths[VG_MAX_OUT], off = 0;
	for (unsigned j = 0; j < P.nout; j++) {
		widths[j] = P.out[j].width;
		cols[j] = b.out + off;
		off += align256(b.rows * widths[j]);
		*d2h += n * widths[j];
	}
	HCHK(hipStreamWaitEvent(down, b.done, 0), "stream order");
	if (n && sink.take(cols, widths, P.nout, n, down) == VIRG_FAIL)
		return VIRG_FAIL;
	HCHK(hipEventRecord(b.freed, down), "event");
	*rows_out += n;
	b.busy = false;
	return VIRG_SUCCESS;
}

static int stream_select(vgd_device *d, const vg_prog &P, const std::vector<const virg_tablet_meta *> &tabs,
	uint64_t batch_rows, stream_sink &sink, vg_stream_stats *st)
{
	const uint32_t T = vg_kernel_tile_rows();
	size_t row_in = 0, row_out = 0;
	for (unsigned c = 0; c < P.ncol; c++)
		row_in += P.col[c].width;
	for (unsigned j = 0; j < P.nout; j++)
		row_out += P.out[j].width;
	/* batches: whole tablets, about batch_rows rows each */
	std::vector<size_t> cut{0};
	uint64_t acc = 0, maxrows = 0, maxtiles = 0, maxtabs = 0;
	for (size_t i = 0; i < tabs.size(); i++) {
		acc += tabs[i]->rows;
		if (acc >= batch_rows || i + 1 == tabs.size()) {
			cut.push_back(i + 1);
			acc = 0;
		}
	}
	for (size_t k = 0; k + 1 < cut.size(); k++) {
		uint64_t r = 0, t = 0;
		for (size_t i = cut[k]; i < cut[k + 1]; i++) {
			r += tabs[i]->rows;
			t += (tabs[i]->rows + T - 1) / T;
		}
		maxrows = r > maxrows ? r : maxrows;
		maxtiles = t > maxtiles ? t : maxtiles;
		maxtabs = cut[k + 1] - cut[k] > maxtabs ? cut[k + 1] - cut[k] : maxtabs;
	}
	/* Tablets the system would not pin are copied by host threads into a
	 * pinned staging buffer first (a whole batch at once). */
	std::vector<uint8_t> pageable(tabs.size());
	bool any_pageable = false;
	for (size_t i = 0; i < tabs.size(); i++) {
		pageable[i] = !host_pinned(tabs[i]);
		any_pageable = any_pageable || pageable[i];
	}
	int rc = VIRG_SUCCESS;
	uint64_t rows_out = 0, h2d = 0, d2h = 0, scanned = 0;
	auto t0 = std::chrono::steady_clock::now();
	if (!d->scache)
		d->scache = new stream_cache();
	sbuf *const B = d->scache->B;
	hipStream_t &down = d->scache->down;
#define SCHK(call, what) do { if (herr((call), (what))) { rc = VIRG_FAIL; goto done; } } while (0)
	if (!down)
		SCHK(hipStreamCreateWithFlags(&down, hipStreamNonBlocking), "stream");
	for (int k = 0; k < 2; k++) {
		sbuf &b = B[k];
		/* staging: every loaded column of every tablet, 256-byte aligned */
		const size_t stage = (size_t)maxrows * row_in + (size_t)maxtabs * P.ncol * 256 + 256;
		const size_t outb = (size_t)maxrows * row_out + 256 * (P.nout + 1);
		b.busy = false;
		if (!b.stream) {
			SCHK(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking), "stream");
			SCHK(hipEventCreateWithFlags(&b.done, hipEventDisableTiming), "event");
			SCHK(hipEventCreateWithFlags(&b.freed, hipEventDisableTiming), "event");
			SCHK(hipHostMalloc((void **)&b.h_ctrl, sizeof(vg_ctrl), hipHostMallocDefault), "control block");
		}
		/* the buffers of the last call, grown where this one needs more */
		if (b.stage_cap < stage) {
			if (b.stage)
				(void)hipFree(b.stage);
			b.stage = nullptr;
			b.stage_cap = 0;
			SCHK(hipMalloc((void **)&b.stage, stage), "stream staging");
			b.stage_cap = stage;
		}
		if (any_pageable && b.h_stage_cap < stage) {
			if (b.h_stage)
				(void)hipHostFree(b.h_stage);
			b.h_stage = nullptr;
			b.h_stage_cap = 0;
			SCHK(hipHostMalloc((void **)&b.h_stage, stage, hipHostMallocDefault),
				"pinned staging for pageable tablets");
			b.h_stage_cap = stage;
		}
		if (b.out_cap < outb) {
			if (b.out)
				(void)hipFree(b.out);
			b.out = nullptr;
			b.out_cap = 0;
			SCHK(hipMalloc((void **)&b.out, outb), "stream output");
			b.out_cap = outb;
		}
		if (b.desc_cap < maxtabs) {
			if (b.d_desc)
				(void)hipFree(b.d_desc);
			if (b.h_desc)
				(void)hipHostFree(b.h_desc);
			b.d_desc = b.h_desc = nullptr;
			b.desc_cap = 0;
			SCHK(hipMalloc((void **)&b.d_desc, maxtabs * sizeof(vg_tabdesc)), "descriptors");
			SCHK(hipHostMalloc((void **)&b.h_desc, maxtabs * sizeof(vg_tabdesc), hipHostMallocDefault),
				"descriptors");
			b.desc_cap = maxtabs;
		}
		if (b.tile_cap < maxtiles) {
			if (b.d_tile)
				(void)hipFree(b.d_tile);
			if (b.h_tile)
				(void)hipHostFree(b.h_tile);
			b.d_tile = b.h_tile = nullptr;
			b.tile_cap = 0;
			SCHK(hipMalloc((void **)&b.d_tile, maxtiles * sizeof(uint32_t)), "tile table");
			SCHK(hipHostMalloc((void **)&b.h_tile, maxtiles * sizeof(uint32_t), hipHostMallocDefault),
				"tile table");
			b.tile_cap = maxtiles;
		}
		if (ensure_lr(b.lr, (uint32_t)maxtiles, vg_kernel_grid(&P, (uint32_t)maxtiles), b.stream) == VIRG_FAIL) {
			rc = VIRG_FAIL;
			goto done;
		}
		SCHK(hipEventRecord(b.freed, down), "event");
	}
	for (size_t k = 0; k + 1 < cut.size(); k++) {
		sbuf &b = B[k & 1];
		if (b.busy && stream_finish(d, P, b, sink, down, &rows_out, &d2h) == VIRG_FAIL) {
			rc = VIRG_FAIL;
			goto done;
		}
		/* stage the batch's loaded columns and describe them */
		size_t off = 0;
		uint32_t ntile = 0;
		plan p;
		b.rows = 0;
		bool staged = false;
		for (size_t i = cut[k]; i < cut[k + 1]; i++)
			staged = staged || pageable[i];
		std::vector<host_copy> jobs;
		for (size_t i = cut[k]; i < cut[k + 1]; i++) {
			const virg_tablet_meta *t = tabs[i];
			vg_tabdesc &td = b.h_desc[i - cut[k]];
			memset(&td, 0, sizeof td);
			for (unsigned c = 0; c < P.ncol; c++) {
				const vg_col &col = P.col[c];
				const size_t bytes = (size_t)t->rows * col.width;
				td.col[col.src] = (uint64_t)(b.stage + off);
				if (staged)
					jobs.push_back({b.h_stage + off, host_column(t, col.src), bytes});
				else if (bytes)
					SCHK(hipMemcpyAsync(b.stage + off, host_column(t, col.src), bytes,
						hipMemcpyHostToDevice, b.stream), "stream columns");
				h2d += bytes;
				off += align256(bytes);
			}
			td.rows = t->rows;
			td.first_tile = ntile;
			td.aligned = 1;
			const uint32_t nt = (t->rows + T - 1) / T;
			for (uint32_t q = 0; q < nt; q++)
				b.h_tile[ntile + q] = (uint32_t)(i - cut[k]);
			ntile += nt;
			b.rows += t->rows;
		}
		scanned += b.rows;
		if (staged) {
			/* the buffer's previous batch has finished (stream_finish), so
			 * its staging copy is no longer being read */
			parallel_copy(jobs);
			SCHK(hipMemcpyAsync(b.stage, b.h_stage, off, hipMemcpyHostToDevice, b.stream),
				"stream columns");
		}
		SCHK(hipMemcpyAsync(b.d_desc, b.h_desc, (cut[k + 1] - cut[k]) * sizeof(vg_tabdesc),
			hipMemcpyHostToDevice, b.stream), "descriptors");
		if (ntile)
			SCHK(hipMemcpyAsync(b.d_tile, b.h_tile, ntile * sizeof(uint32_t),
				hipMemcpyHostToDevice, b.stream), "tile table");
		/* the output buffer is free once the earlier results left it */
		SCHK(hipStreamWaitEvent(b.stream, b.freed, 0), "stream order");
		p.d_desc = b.d_desc;
		p.d_tile = b.d_tile;
		p.num_tiles = ntile;
		p.rows = b.rows;
		{
			void *out[VG_MAX_OUT];
			size_t o = 0;
			for (unsigned j = 0; j < P.nout; j++) {
				out[j] = b.out + o;
				o += align256(b.rows * P.out[j].width);
			}
			b.h_ctrl->total = 0;
			b.h_ctrl->overflow = 0;
			const int lrc = launch(P, p, b.lr, out, b.rows, b.stream);
			p.d_desc = nullptr;     /* owned by the buffer, not the plan */
			p.d_tile = nullptr;
			if (lrc == VIRG_FAIL) {
				rc = VIRG_FAIL;
				goto done;
			}
		}
		if (ntile)
			SCHK(hipMemcpyAsync(b.h_ctrl, b.lr.ctrl, sizeof(vg_ctrl), hipMemcpyDeviceToHost, b.stream),
				"control block");
		SCHK(hipEventRecord(b.done, b.stream), "event");
		b.busy = true;
	}
	for (size_t k = cut.size() - 1; k < cut.size() + 1; k++) {
		sbuf &b = B[k & 1];
		if (b.busy && stream_finish(d, P, b, sink, down, &rows_out, &d2h) == VIRG_FAIL) {
			rc = VIRG_FAIL;
			goto done;
		}
	}
	SCHK(hipStreamSynchronize(down), "result copies");
	if (st) {
		st->rows_scanned = scanned;
		st->rows_out = rows_out;
		st->h2d_bytes = h2d;
		st->d2h_bytes = d2h;
		st->batches = (unsigned)(cut.size() - 1);
		st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	}
done:
#undef SCHK
	/* the buffers stay for the next call; after a failure they go, so that
	 * the next call starts from nothing */
	if (rc == VIRG_FAIL)
		free_stream_cache(d);
	return rc;
}

/* results into caller arrays (pinned host memory copies fastest) */
struct array_sink : stream_sink {
	void **dst;
	uint64_t at = 0, cap;
	array_sink(void **d, uint64_t c) : dst(d), cap(c) {}
	int take(void *const *cols, const size_t *widths, unsigned ncol, uint64_t n, hipStream_t down) override
	{
		VIRG_CHECK(at + n > cap, "stream result exceeds the output capacity");
		for (unsigned j = 0; j < ncol; j++)
			if (dst[j])
				HCHK(hipMemcpyAsync((char *)dst[j] + at * widths[j], cols[j], n * widths[j],
					hipMemcpyDeviceToHost, down), "stream results");
		at += n;
		return VIRG_SUCCESS;
	}
};

/* largest result tablet the device path allocates (rows) */
#define VG_RESULT_TABLET_ROWS (1u << 26)

/* results into the host result tablet chain of a virg_vm (as the reference
 * fills result tablets, allocating the next one when one is full) */
struct tablet_sink : stream_sink {
	virginian *v;
	virg_vm *vm;
	virg_tablet_meta **res;
	tablet_sink(virginian *v_, virg_vm *vm_, virg_tablet_meta **r) : v(v_), vm(vm_), res(r) {}
	int take(void *const *cols, const size_t *widths, unsigned ncol, uint64_t n, hipStream_t down) override
	{
		uint64_t done = 0;
		while (done < n) {
			virg_tablet_meta *r = res[0];
			const uint64_t room = r->possible_rows > r->rows ? r->possible_rows - r->rows : 0;
			const uint64_t k = n - done < room ? n - done : room;
			for (unsigned j = 0; j < ncol && k > 0; j++) {
				char *dst = (char *)r + r->fixed_block + r->fixed_offset[j] + (size_t)r->rows * widths[j];
				HCHK(hipMemcpyAsync(dst, (const char *)cols[j] + done * widths[j], k * widths[j],
					hipMemcpyDeviceToHost, down), "result copy");
			}
			r->rows += (unsigned)k;
			done += k;
			if (done >= n)
				break;
			VIRG_CHECK(r->possible_rows == 0, "result tablet has no capacity");
			virg_tablet_meta *old = r;
			/* the rest of the batch in one tablet (few large copies instead
			 * of one per 8 MB), unless a database file must be able to hold
			 * every tablet */
			const uint64_t left = n - done;
			const unsigned want = (unsigned)(left < VG_RESULT_TABLET_ROWS ? left : VG_RESULT_TABLET_ROWS);
			VIRG_CHECK((v->dbfd < 0 && want > old->possible_rows
					? virg_vm_allocresult_rows(v, vm, res, old, want)
					: virg_vm_allocresult(v, vm, res, old)) == VIRG_FAIL, "result allocation failed");
			virg_tablet_unlock(v, old->id);
		}
		return VIRG_SUCCESS;
	}
};

/* Out-of-core select over the table named by the program's Table op. */
extern "C" int vg_stream_select(virginian *v, const virg_op *ops, unsigned num_ops, void **out,
	unsigned long long capacity, unsigned long long batch_rows, vg_stream_stats *st, char *err,
	size_t errlen)
{
	vgd_device *d = dev_of(v);
	if (!d) {
		snprintf(err, errlen, "no usable GPU: the data-parallel VM needs a HIP device");
		return VIRG_FAIL;
	}
	HCHK(hipSetDevice(d->device), "set device");
	int table = -1, par = -1;
	for (unsigned i = 0; i < num_ops; i++) {
		if (ops[i].op == OP_Table && table < 0)
			table = ops[i].p1;
		if (ops[i].op == OP_Parallel && par < 0)
			par = (int)i;
	}
	if (table < 0 || table >= VIRG_MAX_TABLES || !v->db.table_status[table] || par < 0) {
		snprintf(err, errlen, "program has no Table/Parallel op or names an unknown table");
		return VIRG_FAIL;
	}
	virg_tablet_meta *t;
	if (virg_db_load(v, v->db.first_tablet[table], &t) == VIRG_FAIL) {
		snprintf(err, errlen, "table has no tablets");
		return VIRG_FAIL;
	}
	vg_prog P;
	if (vg_decode(ops, num_ops, (unsigned)par + 1, t, nullptr, &P, err, errlen) == VIRG_FAIL ||
		vg_kernel_check(&P) != 0) {
		if (!err[0])
			snprintf(err, errlen, "op-program needs more register slots than the LDS holds");
		virg_tablet_unlock(v, t->id);
		return VIRG_FAIL;
	}
	/* the tablets' host memory is read while streaming: with a database
	 * file open each keeps a lock until then (the slot pager) */
	std::vector<const virg_tablet_meta *> tabs;
	held_locks held(v);
	while (1) {
		tabs.push_back(t);
		if (v->dbfd >= 0 && held.take(t->id) == VIRG_FAIL) {
			virg_tablet_unlock(v, t->id);
			snprintf(err, errlen, "could not lock a tablet");
			return VIRG_FAIL;
		}
		if (t->last_tablet)
			break;
		if (virg_db_loadnext(v, &t) == VIRG_FAIL) {
			virg_tablet_unlock(v, t->id);
			snprintf(err, errlen, "could not load the next tablet: every tablet in host memory is "
				"locked (a streamed table keeps each locked; raise vg_tablet_limit)");
			return VIRG_FAIL;
		}
	}
	virg_tablet_unlock(v, t->id);
	array_sink sink(out, capacity);
	if (stream_select(d, P, tabs, batch_rows ? batch_rows : (8u << 20), sink, st) == VIRG_FAIL) {
		snprintf(err, errlen, "stream failed (see stderr)");
		return VIRG_FAIL;
	}
	return VIRG_SUCCESS;
}

/* Pinned host memory, for stream outputs that copy at link speed. */
extern "C" void *vg_host_alloc(size_t bytes)
{
	void *p = nullptr;
	if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
		(void)hipGetLastError();
		return nullptr;
	}
	return p;
}

extern "C" void vg_host_free(void *p)
{
	if (p)
		(void)hipHostFree(p);
}

/* Host-to-device copy rate of pinned memory (one copy of `bytes`, best of
 * `reps`), for comparing the stream's link rate with what the link gives. */
extern "C" int vg_link_probe(virginian *v, size_t bytes, int reps, double *h2d_gbs, double *d2h_gbs)
{
	vgd_device *d = dev_of(v);
	VIRG_CHECK(!d, "no device");
	HCHK(hipSetDevice(d->device), "set device");
	void *h = nullptr, *g = nullptr;
	HCHK(hipHostMalloc(&h, bytes, hipHostMallocDefault), "pinned buffer");
	if (hipMalloc(&g, bytes) != hipSuccess) {
		(void)hipHostFree(h);
		return VIRG_FAIL;
	}
	memset(h, 1, bytes);
	hipEvent_t a, b;
	(void)hipEventCreate(&a);
	(void)hipEventCreate(&b);
	float best[2] = {1e30f, 1e30f};
	for (int dir = 0; dir < 2; dir++)
		for (int i = 0; i < reps; i++) {
			(void)hipEventRecord(a, d->stream);
			if (dir == 0)
				(void)hipMemcpyAsync(g, h, bytes, hipMemcpyHostToDevice, d->stream);
			else
				(void)hipMemcpyAsync(h, g, bytes, hipMemcpyDeviceToHost, d->stream);
			(void)hipEventRecord(b, d->stream);
			(void)hipEventSynchronize(b);
			float ms = 0;
			(void)hipEventElapsedTime(&ms, a, b);
			if (ms < best[dir])
				best[dir] = ms;
		}
	(void)hipEventDestroy(a);
	(void)hipEventDestroy(b);
	(void)hipFree(g);
	(void)hipHostFree(h);
	h2d_gbs[0] = (double)bytes / (best[0] * 1e6);
	d2h_gbs[0] = (double)bytes / (best[1] * 1e6);
	return VIRG_SUCCESS;
}

/* ------------------------------------------------------------ boundary */

/* reference: src/vm/vm_gpu.cu:27 */
extern "C" const size_t *virg_gpu_getsizes(void)
{
	return &virg_testsizes[0];
}

/* reference: src/vm/vm_gpu.cu:107 */

extern "C" int virg_vm_gpu(virginian *v, virg_vm *vm, virg_tablet_meta **tab,
	virg_tablet_meta **res, unsigned num_tablets)
{
	vgd_device *d = dev_of(v);
	VIRG_CHECK(d == nullptr, "no usable GPU: the data-parallel VM needs a HIP device");
	VIRG_CHECK(v->threads_per_block != VIRG_THREADSPERBLOCK, "Cannot change compile-time threads per block");
	HCHK(hipSetDevice(d->device), "set device");

	char err[256];
	vg_prog P;
	if (vg_decode(vm->stmt, vm->num_ops, vm->pc, tab[0], res[0], &P, err, sizeof err) == VIRG_FAIL) {
		fprintf(stderr, "::: Virginian: op-program not supported on the GPU: %s\n", err);
		return VIRG_FAIL;
	}
	VIRG_CHECK(vg_kernel_check(&P) != 0, "op-program needs more register slots than the LDS holds");

	/* collect the tablets: tab[0] arrives locked; like the reference the
	 * walk unlocks each tablet as it moves on and leaves the last locked.
	 * With a database file open an unlocked tablet may be written out and
	 * its slot reused (the slot pager), so the resident path makes each
	 * tablet resident in HBM while it is locked and keeps a copy of its
	 * meta, and the streaming path, which reads host memory later, keeps a
	 * lock on every tablet until it is done. */
	const bool paged = v->dbfd >= 0;
	/* The reference's mapped mode (use_mmap, src/vm/vm_gpu.cu:445-560) runs
	 * each tablet where it lies in host memory and writes the results into
	 * the host result tablets.  Here both host-memory modes take the
	 * streaming path: the data stays in host memory and crosses the link
	 * once, by DMA copies of only the referenced columns, which reach the
	 * link's rate where compute-unit loads of mapped memory would not;
	 * results land in the host result tablets as in the reference. */
	const bool host_data = v->use_stream || v->use_mmap;
	std::vector<const virg_tablet_meta *> tabs;
	std::vector<virg_tablet_meta> metas;
	held_locks held(v);
	unsigned proced = 0;
	while (1) {
		if (paged && !host_data) {
			VIRG_CHECK(!resident_tablet(d, tab[0], d->stream) || hipStreamSynchronize(d->stream) != hipSuccess,
				"could not make a tablet resident");
			metas.push_back(*tab[0]);
		} else {
			tabs.push_back(tab[0]);
			if (paged)
				VIRG_CHECK(held.take(tab[0]->id) == VIRG_FAIL, "could not lock a tablet");
		}
		proced++;
		if (tab[0]->last_tablet || (num_tablets != 0 && proced >= num_tablets))
			break;
		VIRG_CHECK(virg_db_loadnext(v, tab) == VIRG_FAIL, "could not load next tablet");
		if (tab[0]->rows == 0)
			break;
	}
	for (const virg_tablet_meta &m : metas)
		tabs.push_back(&m);

	if (host_data) {
		/* the reference's streaming and mapped modes: the tablets stay in
		 * host memory and stream through HBM batch by batch */
		tablet_sink sink(v, vm, res);
		vg_stream_stats st;
		memset(&st, 0, sizeof st);
		if (stream_select(d, P, tabs, 8u << 20, sink, &st) == VIRG_FAIL)
			return VIRG_FAIL;
		vm->timing1 = 0.0f;
		vm->timing2 = (float)st.seconds;
		vm->timing3 = 0.0f;
		return VIRG_SUCCESS;
	}

	/* timing events, destroyed on every way out */
	struct events {
		hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
		~events()
		{
			for (hipEvent_t x : e)
				if (x)
					(void)hipEventDestroy(x);
		}
	} evs;
	hipEvent_t *ev = evs.e;
	for (int i = 0; i < 4; i++)
		HCHK(hipEventCreate(&ev[i]), "event");
	HCHK(hipEventRecord(ev[0], d->stream), "event");
	plan p;
	if (build_plan(d, tabs, d->stream, p) == VIRG_FAIL) {
		free_plan(p);
		return VIRG_FAIL;
	}
	HCHK(hipEventRecord(ev[1], d->stream), "event");

	/* device output columns, capacity = every row selected */
	size_t widths[VG_MAX_OUT], total_w = 0;
	for (unsigned j = 0; j < P.nout; j++) {
		widths[j] = P.out[j].width;
		total_w += widths[j];
	}
	size_t need = (size_t)p.rows * total_w + 256 * (P.nout + 1);
	if (need > d->outbuf_bytes) {
		/* grown at least twofold, so a run of ever larger results
		 * reallocates (and waits for the device to free) only a few times */
		const size_t want = need > 2 * d->outbuf_bytes ? need : 2 * d->outbuf_bytes;
		if (d->outbuf) {
			HCHK(hipStreamSynchronize(d->stream), "sync");
			HCHK(hipFree(d->outbuf), "free");
		}
		d->outbuf = nullptr;
		d->outbuf_bytes = 0;
		HCHK(hipMalloc(&d->outbuf, want), "result buffer");
		d->outbuf_bytes = want;
	}
	void *out[VG_MAX_OUT];
	size_t off = 0;
	for (unsigned j = 0; j < P.nout; j++) {
		out[j] = (char *)d->outbuf + off;
		off += ((size_t)p.rows * widths[j] + 255) / 256 * 256;
	}

	int ok = launch(P, p, d->lr, out, p.rows, d->stream);
	HCHK(hipEventRecord(ev[2], d->stream), "event");
	uint64_t nres = 0;
	if (ok == VIRG_SUCCESS && p.num_tiles)
		ok = read_ctrl(d->lr, d->stream, &nres);
	free_plan(p);
	if (ok == VIRG_FAIL)
		return VIRG_FAIL;

	/* fill host result tablets in order */
	uint64_t done = 0;
	while (1) {
		virg_tablet_meta *r = res[0];
		uint64_t room = r->possible_rows > r->rows ? r->possible_rows - r->rows : 0;
		uint64_t n = nres - done < room ? nres - done : room;
		for (unsigned j = 0; j < P.nout && n > 0; j++) {
			char *dst = (char *)r + r->fixed_block + r->fixed_offset[j] + (size_t)r->rows * widths[j];
			HCHK(hipMemcpyAsync(dst, (char *)out[j] + done * widths[j], n * widths[j],
				hipMemcpyDeviceToHost, d->stream), "result copy");
		}
		r->rows += (unsigned)n;
		done += n;
		if (done >= nres)
			break;
		VIRG_CHECK(r->possible_rows == 0, "result tablet has no capacity");
		virg_tablet_meta *old = r;
		/* the rest in one tablet (few large copies instead of one per 8 MB),
		 * unless a database file must be able to hold every tablet */
		const uint64_t left = nres - done;
		const unsigned want = (unsigned)(left < VG_RESULT_TABLET_ROWS ? left : VG_RESULT_TABLET_ROWS);
		VIRG_CHECK((v->dbfd < 0 && want > old->possible_rows
				? virg_vm_allocresult_rows(v, vm, res, old, want)
				: virg_vm_allocresult(v, vm, res, old)) == VIRG_FAIL, "result allocation failed");
		virg_tablet_unlock(v, old->id);
	}
	HCHK(hipEventRecord(ev[3], d->stream), "event");
	HCHK(hipStreamSynchronize(d->stream), "result copy");
	float ms[3];
	for (int i = 0; i < 3; i++)
		HCHK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]), "event time");
	vm->timing1 = ms[0] / 1000.0f;
	vm->timing2 = ms[1] / 1000.0f;
	vm->timing3 = ms[2] / 1000.0f;
	return VIRG_SUCCESS;
}

/* The shared plan of a table's current tablets, built on first use. */
static plan *shared_plan_of(vgd_device *d, unsigned table, const std::vector<const virg_tablet_meta *> &tabs)
{
	std::vector<std::pair<unsigned, unsigned>> key;
	for (const virg_tablet_meta *t : tabs)
		key.push_back({t->id, t->rows});
	auto it = d->plans.find(table);
	if (it != d->plans.end()) {
		if (it->second->key == key)
			return &it->second->p;
		d->retired.push_back(it->second);
		d->plans.erase(it);
	}
	shared_plan *sp = new shared_plan();
	sp->key = key;
	if (build_plan(d, tabs, d->stream, sp->p) == VIRG_FAIL || hipStreamSynchronize(d->stream) != hipSuccess) {
		free_plan(sp->p);
		delete sp;
		return nullptr;
	}
	d->plans[table] = sp;
	return &sp->p;
}

/* a control block from the device's pool */
static vg_ctrl *pool_ctrl(vgd_device *d)
{
	if (d->ctrl_free.empty()) {
		const unsigned chunk = 256;
		vg_ctrl *c = nullptr;
		if (hipMalloc((void **)&c, chunk * sizeof(vg_ctrl)) != hipSuccess) {
			(void)hipGetLastError();
			return nullptr;
		}
		if (hipMemset(c, 0, chunk * sizeof(vg_ctrl)) != hipSuccess) {
			(void)hipGetLastError();
			(void)hipFree(c);
			return nullptr;
		}
		d->ctrl_chunks.push_back(c);
		for (unsigned i = 0; i < chunk; i++)
			d->ctrl_free.push_back(c + chunk - 1 - i);
	}
	/* no clearing: a batch run's reset and a two-pass launch's count pass
	 * clear what they use */
	vg_ctrl *c = d->ctrl_free.back();
	d->ctrl_free.pop_back();
	return c;
}

/* A light query's two-pass arrays: pooled ones large enough, if any. */
static void take_pooled_lr(vgd_device *d, launch_res &lr, uint32_t num_tiles, uint32_t grid)
{
	if (lr.vbits)
		return;
	for (size_t i = 0; i < d->lr_free.size(); i++) {
		launch_res &f = d->lr_free[i];
		if (f.tiles_cap >= num_tiles && f.wgcount_cap >= grid) {
			vg_ctrl *c = lr.ctrl;
			lr = f;
			lr.ctrl = c;
			d->lr_free.erase(d->lr_free.begin() + (long)i);
			return;
		}
	}
}

/* -------------------------------------------------------- vg_query ABI */
struct vg_query;
static int ensure_events(vg_query *q);


struct vg_query {
	virginian *v = nullptr;
	vgd_device *d = nullptr;
	vg_prog P;
	std::vector<const virg_tablet_meta *> tabs;
	std::vector<virg_tablet_meta> tabmeta;   /* copies of the metas, when the tablets may be paged out */
	plan own;
	plan *p = &own;          /* own, or the device's shared plan of the table */
	bool own_stream = true;  /* else the device's stream */
	bool own_out = true;     /* else the outputs are the caller's memory */
	bool pooled_ctrl = false;
	launch_res lr;
	void *out[VG_MAX_OUT];
	size_t width[VG_MAX_OUT];
	uint64_t capacity = 0;
	hipStream_t stream = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	uint64_t last_rows = 0;
	uint64_t read_bytes_per_row = 0;
	/* the program's specialised single-pass kernel (jit.hip), if it has one */
	vgx_kernel *jk = nullptr;
	vgx_tile *d_jtiles = nullptr;
	uint64_t *d_jstate = nullptr;      /* chunk counter, epoch, one status word per chunk */
	uint32_t jgrid = 0;
	vgx_launch_args ja;
	/* batches (vg_batch, vg_pbatch) holding this query's launch arguments:
	 * while any does, its outputs may not change and it may not be freed */
	unsigned batch_refs = 0;
};

static void refs_take(const std::vector<vg_query *> &qs)
{
	for (vg_query *q : qs)
		q->batch_refs++;
}

static void refs_drop(const std::vector<vg_query *> &qs)
{
	for (vg_query *q : qs)
		if (q->batch_refs)
			q->batch_refs--;
}

/* timing events, made on first use (light queries have none) */
static int ensure_events(vg_query *q)
{
	if (!q->e0)
		HCHK(hipEventCreate(&q->e0), "event");
	if (!q->e1)
		HCHK(hipEventCreate(&q->e1), "event");
	return VIRG_SUCCESS;
}

static int qerr(char *err, size_t n, const char *fmt, ...)
{
	if (err && n) {
		va_list ap;
		va_start(ap, fmt);
		vsnprintf(err, n, fmt, ap);
		va_end(ap);
	}
	return VIRG_FAIL;
}

/* When prepared queries get a specialised kernel: 0 never (the two-pass
 * interpreter only), 1 queries prepared with vg_query_create other than
 * light ones and those flagged VG_QUERY_INTERP, 2 also light queries and
 * the drop-in path.  The environment variable VG_JIT sets the initial
 * value (default 1). */
static int g_jit_policy = -1;

static int jit_policy()
{
	if (g_jit_policy < 0) {
		const char *e = getenv("VG_JIT");
		g_jit_policy = e && *e ? atoi(e) : 1;
	}
	return g_jit_policy;
}

extern "C" int vg_jit_policy(int policy)
{
	const int old = jit_policy();
	if (policy >= 0)
		g_jit_policy = policy;
	return old;
}

/* Tile table of a plan for tiles of TR rows: (tablet, first row) of each. */
static std::vector<vgx_tile> jit_tiles(const plan &p, uint32_t TR)
{
	std::vector<vgx_tile> t;
	for (size_t i = 0; i < p.desc.size(); i++)
		for (uint32_t r = 0; r < p.desc[i].rows; r += TR)
			t.push_back(vgx_tile{(uint32_t)i, r});
	return t;
}

/* Give q its specialised kernel, tile table and launch state.  Returns 0,
 * or -1 (with a message) when the program has no specialised form; the
 * query then runs on the interpreter. */
static int jit_setup(vg_query *q, char *err, size_t errlen)
{
	const plan &p = *q->p;
	bool aligned = true;
	for (const vg_tabdesc &d : p.desc)
		aligned = aligned && d.aligned;
	vgx_kernel *k = vgx_get(&q->P, aligned ? 1 : 0, err, errlen);
	if (!k)
		return -1;
	const vgx_shape *sh = vgx_kernel_shape(k);
	const std::vector<vgx_tile> tiles = jit_tiles(p, sh->TR);
	if (tiles.empty())
		return -1;
	const uint32_t nchunks = (uint32_t)((tiles.size() + sh->S - 1) / sh->S);
	const size_t state_bytes = (2 + (size_t)nchunks) * sizeof(uint64_t);
	if (hipMalloc((void **)&q->d_jtiles, tiles.size() * sizeof(vgx_tile)) != hipSuccess ||
		hipMalloc((void **)&q->d_jstate, state_bytes) != hipSuccess ||
		hipMemcpyAsync(q->d_jtiles, tiles.data(), tiles.size() * sizeof(vgx_tile), hipMemcpyHostToDevice,
			q->stream) != hipSuccess ||
		hipMemsetAsync(q->d_jstate, 0, state_bytes, q->stream) != hipSuccess ||
		hipStreamSynchronize(q->stream) != hipSuccess) {
		(void)hipGetLastError();
		qerr(err, errlen, "cannot set up the specialised kernel's launch state");
		return -1;
	}
	vgx_launch_args &a = q->ja;
	memset(&a, 0, sizeof a);
	a.tabs = p.d_desc;
	a.tiles = q->d_jtiles;
	a.num_tiles = (uint32_t)tiles.size();
	a.nchunks = nchunks;
	a.state = q->d_jstate;
	a.status = q->d_jstate + 2;
	a.ctrl = q->lr.ctrl;
	a.capacity = q->capacity;
	a.spin_limit = 1u << 20;
	for (unsigned j = 0; j < q->P.nout; j++)
		a.out[j] = q->out[j];
	vgx_immediates(&q->P, &a);
	q->jgrid = vgx_kernel_grid(k, nchunks, (unsigned)q->d->prop.multiProcessorCount);
	q->jk = k;
	return 0;
}

/* One run of q on `stream`: its specialised kernel, or the interpreter's
 * count and write passes. */
static int query_launch(vg_query *q, hipStream_t stream, hipEvent_t mid = nullptr)
{
	if (q->jk) {
		if (mid)
			HCHK(hipEventRecord(mid, stream), "event");
		if (vgx_launch(q->jk, &q->ja, q->jgrid, stream) != 0) {
			herr(hipGetLastError(), "specialised kernel launch");
			return VIRG_FAIL;
		}
		return VIRG_SUCCESS;
	}
	return launch(q->P, *q->p, q->lr, q->out, q->capacity, stream, mid);
}

/* For the serving loop: run q alone on `stream` (its count and write
 * passes), the write pass storing the outcome into res (pinned host
 * memory, cleared here) once every result row is in place.  Interpreter
 * queries only. */
extern "C" int vgd_query_launch_lone(vg_query *q, hipStream_t stream, vg_bresult *res)
{
	if (q->jk || q->batch_refs)
		return VIRG_FAIL;
	vg_bresult *d_res = nullptr;
	HCHK(hipHostGetDevicePointer((void **)&d_res, res, 0), "outcome");
	memset(res, 0, sizeof *res);
	if (q->pooled_ctrl)
		take_pooled_lr(q->d, q->lr, q->p->num_tiles, vg_kernel_grid(&q->P, q->p->num_tiles ? q->p->num_tiles : 1));
	if (q->p->num_tiles == 0) {
		res->done = 1;
		return VIRG_SUCCESS;
	}
	return launch(q->P, *q->p, q->lr, q->out, q->capacity, stream, nullptr, d_res);
}

extern "C" void vg_query_free(vg_query *q)
{
	if (!q)
		return;
	if (q->batch_refs) {
		fprintf(stderr, "::: Virginian: vg_query_free refused: %u batch(es) still hold the query "
			"(free or reload them first)\n", q->batch_refs);
		return;
	}
	/* a light query is freed after its results were collected */
	if (q->stream && !q->pooled_ctrl)
		(void)hipStreamSynchronize(q->stream);
	free_plan(q->own);
	if (q->d_jtiles)
		(void)hipFree(q->d_jtiles);
	if (q->d_jstate)
		(void)hipFree(q->d_jstate);
	if (q->pooled_ctrl && q->lr.ctrl) {
		q->d->ctrl_free.push_back(q->lr.ctrl);
		q->lr.ctrl = nullptr;
		if (q->lr.vbits) {
			/* keep the two-pass arrays for the next light query (a free
			 * would wait for the whole device) */
			q->d->lr_free.push_back(q->lr);
			q->lr = launch_res();
		}
	}
	free_lr(q->lr);
	for (unsigned j = 0; j < VG_MAX_OUT && q->own_out; j++)
		if (q->out[j])
			(void)hipFree(q->out[j]);
	if (q->e0) (void)hipEventDestroy(q->e0);
	if (q->e1) (void)hipEventDestroy(q->e1);
	if (q->stream && q->own_stream) (void)hipStreamDestroy(q->stream);
	delete q;
}

/* Prepare the data-parallel section of a generated program (ops from the
 * SQL front-end or virg_vm_addop: Table, ResultColumn..., Parallel, ...,
 * Result, Converge, Finish) for repeated execution over its table with the
 * table resident in HBM.  Output capacity is `capacity` rows (0 = every
 * row of the table).  `reserved` must be 0. */
extern "C" vg_query *vg_query_create(virginian *v, const virg_op *ops, unsigned num_ops,
	unsigned long long capacity, int flags, char *err, size_t errlen)
{
	vgd_device *d = dev_of(v);
	if (!d) {
		qerr(err, errlen, "no usable GPU: the data-parallel VM needs a HIP device");
		return nullptr;
	}
	if (hipSetDevice(d->device) != hipSuccess) {
		qerr(err, errlen, "cannot select device");
		return nullptr;
	}
	int table = -1, par = -1;
	virg_tablet_meta res;
	memset(&res, 0, sizeof res);
	for (unsigned i = 0; i < num_ops; i++) {
		if (ops[i].op == OP_Table && table < 0)
			table = ops[i].p1;
		if (ops[i].op == OP_ResultColumn && res.fixed_columns < VIRG_MAX_COLUMNS) {
			res.fixed_type[res.fixed_columns] = (virg_t)ops[i].p1;
			res.fixed_stride[res.fixed_columns] = virg_sizeof((virg_t)ops[i].p1);
			res.fixed_columns++;
		}
		if (ops[i].op == OP_Parallel && par < 0)
			par = (int)i;
	}
	if (table < 0 || table >= VIRG_MAX_TABLES || !v->db.table_status[table] || par < 0) {
		qerr(err, errlen, "program has no Table/Parallel op or names an unknown table");
		return nullptr;
	}
	vg_query *q = new vg_query();
	memset(q->out, 0, sizeof q->out);
	q->v = v;
	q->d = d;
	virg_tablet_meta *t;
	if (virg_db_load(v, v->db.first_tablet[table], &t) == VIRG_FAIL) {
		qerr(err, errlen, "table has no tablets");
		delete q;
		return nullptr;
	}
	if (vg_decode(ops, num_ops, (unsigned)par + 1, t, &res, &q->P, err, errlen) == VIRG_FAIL) {
		virg_tablet_unlock(v, t->id);
		delete q;
		return nullptr;
	}
	/* with a database file open an unlocked tablet may be written out and
	 * its slot reused (the slot pager): each tablet is made resident in HBM
	 * while it is locked and the query keeps a copy of its meta */
	const bool paged = v->dbfd >= 0;
	while (1) {
		if (paged) {
			if (!resident_tablet(d, t, d->stream) || hipStreamSynchronize(d->stream) != hipSuccess) {
				virg_tablet_unlock(v, t->id);
				qerr(err, errlen, "cannot make the table resident");
				delete q;
				return nullptr;
			}
			q->tabmeta.push_back(*t);
		} else {
			q->tabs.push_back(t);
		}
		if (t->last_tablet)
			break;
		if (virg_db_loadnext(v, &t) == VIRG_FAIL) {
			virg_tablet_unlock(v, t->id);
			qerr(err, errlen, "cannot load the table's next tablet");
			delete q;
			return nullptr;
		}
	}
	virg_tablet_unlock(v, t->id);
	for (const virg_tablet_meta &m : q->tabmeta)
		q->tabs.push_back(&m);
	if (vg_kernel_check(&q->P) != 0) {
		qerr(err, errlen, "op-program needs more register slots than the LDS holds");
		delete q;
		return nullptr;
	}
	const bool light = (flags & VG_QUERY_LIGHT) != 0;
	if (light) {
		q->own_stream = false;
		q->stream = d->stream;
		q->p = shared_plan_of(d, (unsigned)table, q->tabs);
		if (!q->p) {
			q->p = &q->own;
			qerr(err, errlen, "cannot make the table resident");
			vg_query_free(q);
			return nullptr;
		}
	} else {
		if (hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) != hipSuccess ||
			hipEventCreate(&q->e0) != hipSuccess || hipEventCreate(&q->e1) != hipSuccess) {
			qerr(err, errlen, "cannot create stream/events");
			vg_query_free(q);
			return nullptr;
		}
		if (build_plan(d, q->tabs, q->stream, q->own) == VIRG_FAIL) {
			qerr(err, errlen, "cannot make the table resident");
			vg_query_free(q);
			return nullptr;
		}
	}
	q->capacity = capacity ? capacity : q->p->rows;
	for (unsigned j = 0; j < q->P.nout; j++)
		q->width[j] = q->P.out[j].width;
	if (flags & VG_QUERY_NO_OUTPUT) {
		q->own_out = false;
	} else {
		for (unsigned j = 0; j < q->P.nout; j++)
			if (hipMalloc(&q->out[j], q->capacity * q->width[j] + 256) != hipSuccess) {
				qerr(err, errlen, "cannot allocate %llu-row output", (unsigned long long)q->capacity);
				vg_query_free(q);
				return nullptr;
			}
	}
	uint64_t rb = 0;
	for (unsigned c = 0; c < q->P.ncol; c++)
		rb += q->P.col[c].width;
	q->read_bytes_per_row = rb;
	if (light) {
		/* the control block only; the two-pass arrays on first use */
		q->lr.ctrl = pool_ctrl(d);
		q->pooled_ctrl = q->lr.ctrl != nullptr;
		if (!q->lr.ctrl) {
			qerr(err, errlen, "cannot set up the launch");
			vg_query_free(q);
			return nullptr;
		}
	} else if (ensure_lr(q->lr, q->p->num_tiles, vg_kernel_grid(&q->P, q->p->num_tiles), q->stream) == VIRG_FAIL ||
		hipStreamSynchronize(q->stream) != hipSuccess) {
		qerr(err, errlen, "cannot set up the launch");
		vg_query_free(q);
		return nullptr;
	}
	/* A table of fewer than 8K rows per compute unit gives the single pass
	 * fewer chunks than CUs: its chunk chain cannot fill the device, and
	 * the interpreter's two launches over every CU are as fast (1M rows on
	 * 256 CUs: 17.4 us against 22.8 us at 64 chunks, 17.9 us at 256) with
	 * no compile.  Policy 2 takes every query. */
	const bool small = q->p->rows < (uint64_t)q->d->prop.multiProcessorCount * 8192u;
	const int pol = jit_policy();
	if (q->p->num_tiles && !(flags & VG_QUERY_INTERP) && (pol >= 2 || (pol == 1 && !light && !small))) {
		char jerr[512];
		if (jit_setup(q, jerr, sizeof jerr) != 0) {
			static bool said = false;
			if (!said)
				fprintf(stderr, "::: Virginian: no specialised kernel (%s); the interpreter runs it\n", jerr);
			said = true;
		}
	}
	return q;
}

/* Launch once, asynchronously, on the query's stream. */
extern "C" int vg_query_launch(vg_query *q)
{
	if (q->pooled_ctrl && !q->jk)
		take_pooled_lr(q->d, q->lr, q->p->num_tiles, vg_kernel_grid(&q->P, q->p->num_tiles));
	return query_launch(q, q->stream);
}

/* Wait for the last launch; returns the selected row count. */
extern "C" int vg_query_sync(vg_query *q, unsigned long long *rows)
{
	uint64_t n = 0;
	if (q->jk && q->p->num_tiles) {
		/* a look-back that gave up waiting leaves an unusable result: run
		 * the interpreter's two passes instead */
		vg_ctrl c;
		HCHK(hipMemcpyAsync(&c, q->lr.ctrl, sizeof c, hipMemcpyDeviceToHost, q->stream), "control block");
		HCHK(hipStreamSynchronize(q->stream), "kernel");
		if (c.overflow & VG_OVF_WAIT) {
			fprintf(stderr, "::: Virginian: single-pass wait gave up; rerunning in two passes\n");
			HCHK(hipMemsetAsync(q->lr.ctrl, 0, sizeof(vg_ctrl), q->stream), "control block");
			if (launch(q->P, *q->p, q->lr, q->out, q->capacity, q->stream) == VIRG_FAIL)
				return VIRG_FAIL;
		}
	}
	if (q->p->num_tiles && read_ctrl(q->lr, q->stream, &n) == VIRG_FAIL)
		return VIRG_FAIL;
	q->last_rows = n;
	rows[0] = n;
	return VIRG_SUCCESS;
}

extern "C" int vg_query_run(vg_query *q, unsigned long long *rows)
{
	if (vg_query_launch(q) == VIRG_FAIL)
		return VIRG_FAIL;
	return vg_query_sync(q, rows);
}

/* Copy the first min(rows, max_rows) result rows to host arrays. */
extern "C" int vg_query_fetch(vg_query *q, void **cols, unsigned long long max_rows)
{
	uint64_t n = q->last_rows < max_rows ? q->last_rows : max_rows;
	for (unsigned j = 0; j < q->P.nout; j++)
		if (cols[j] && n)
			HCHK(hipMemcpyAsync(cols[j], q->out[j], n * q->width[j], hipMemcpyDeviceToHost, q->stream), "fetch");
	HCHK(hipStreamSynchronize(q->stream), "fetch");
	return VIRG_SUCCESS;
}

/* Time `iters` back-to-back launches after `warmup` untimed ones with HIP
 * events on the query's stream (the stream the kernels run on).  Reports
 * the mean time per launch pair. */
extern "C" int vg_query_time(vg_query *q, int warmup, int iters, double *ms_per_launch,
	unsigned long long *rows)
{
	for (int i = 0; i < warmup; i++)
		if (vg_query_launch(q) == VIRG_FAIL)
			return VIRG_FAIL;
	if (vg_query_sync(q, rows) == VIRG_FAIL)
		return VIRG_FAIL;
	if (ensure_events(q) == VIRG_FAIL)
		return VIRG_FAIL;
	HCHK(hipEventRecord(q->e0, q->stream), "event");
	for (int i = 0; i < iters; i++)
		if (vg_query_launch(q) == VIRG_FAIL)
			return VIRG_FAIL;
	HCHK(hipEventRecord(q->e1, q->stream), "event");
	if (vg_query_sync(q, rows) == VIRG_FAIL)
		return VIRG_FAIL;
	float ms = 0;
	HCHK(hipEventElapsedTime(&ms, q->e0, q->e1), "event time");
	ms_per_launch[0] = iters > 0 ? ms / iters : 0.0;
	return VIRG_SUCCESS;
}

/* Per-pass times: `iters` launches after `warmup`, each bracketed by
 * events on the query's stream (start, between the count and the write
 * pass, end) and waited for before the next; means in milliseconds. */
extern "C" int vg_query_time_passes(vg_query *q, int warmup, int iters, double *ms_count,
	double *ms_write, unsigned long long *rows)
{
	if (ensure_events(q) == VIRG_FAIL)
		return VIRG_FAIL;
	hipEvent_t mid = nullptr;
	HCHK(hipEventCreate(&mid), "event");
	double c = 0, w = 0;
	int rc = VIRG_SUCCESS;
	for (int i = 0; i < warmup + iters && rc == VIRG_SUCCESS; i++) {
		if (hipEventRecord(q->e0, q->stream) != hipSuccess ||
			query_launch(q, q->stream, mid) == VIRG_FAIL ||
			hipEventRecord(q->e1, q->stream) != hipSuccess ||
			vg_query_sync(q, rows) == VIRG_FAIL) {
			rc = VIRG_FAIL;
			break;
		}
		float a = 0, b = 0;
		if (hipEventElapsedTime(&a, q->e0, mid) != hipSuccess ||
			hipEventElapsedTime(&b, mid, q->e1) != hipSuccess) {
			rc = VIRG_FAIL;
			break;
		}
		if (i >= warmup) {
			c += a;
			w += b;
		}
	}
	(void)hipEventDestroy(mid);
	ms_count[0] = iters > 0 ? c / iters : 0.0;
	ms_write[0] = iters > 0 ? w / iters : 0.0;
	return rc;
}

/* Shape of a prepared query: rows scanned, tiles, referenced-column bytes
 * per row, result bytes per row, result columns. */
extern "C" int vg_query_info(vg_query *q, unsigned long long *rows, unsigned *tiles,
	unsigned *read_bytes_per_row, unsigned *write_bytes_per_row, unsigned *nout)
{
	rows[0] = q->p->rows;
	tiles[0] = q->p->num_tiles;
	read_bytes_per_row[0] = (unsigned)q->read_bytes_per_row;
	unsigned w = 0;
	for (unsigned j = 0; j < q->P.nout; j++)
		w += (unsigned)q->width[j];
	write_bytes_per_row[0] = w;
	nout[0] = q->P.nout;
	return VIRG_SUCCESS;
}

/* What runs q: *fused = 1 for a specialised single-pass kernel (then its
 * rows per lane, tiles per chunk, grid and compile time), 0 for the
 * interpreter's two passes. */
extern "C" int vg_query_kernel(vg_query *q, int *fused, int *rows_per_lane, unsigned *tiles_per_chunk,
	unsigned *grid, double *compile_ms, int *mode)
{
	*fused = q->jk != nullptr;
	/* bit 0: the aligned variant; bit 1: filter columns staged in LDS;
	 * bit 2: the filter column held in registers; bit 3: pipelined; bits
	 * 8-15: the waves per SIMD it was compiled for (0: the compiler's) */
	*mode = 0;
	if (q->jk) {
		const vgx_shape *sh = vgx_kernel_shape(q->jk);
		*mode = (sh->aligned ? 1 : 0) | (sh->staged ? 2 : 0) | (sh->held ? 4 : 0) | (sh->pipe ? 8 : 0) |
			((sh->waves & 0xff) << 8);
	}
	*rows_per_lane = q->jk ? vgx_kernel_shape(q->jk)->R : 0;
	*tiles_per_chunk = q->jk ? vgx_kernel_shape(q->jk)->S : 0;
	*grid = q->jk ? q->jgrid : 0;
	*compile_ms = q->jk ? vgx_kernel_compile_ms(q->jk) : 0.0;
	return VIRG_SUCCESS;
}

/* Name and compute-unit count of the device, 0 if there is none. */
extern "C" int vg_device_info(virginian *v, char *name, size_t n, int *cus)
{
	vgd_device *d = dev_of(v);
	if (!d) {
		if (n) name[0] = '\0';
		cus[0] = 0;
		return VIRG_FAIL;
	}
	snprintf(name, n, "%s", d->prop.gcnArchName);
	cus[0] = d->prop.multiProcessorCount;
	return VIRG_SUCCESS;
}

/* Decode only (no device needed): used to inspect the micro-program. */
extern "C" int vg_decode_program(const virg_op *ops, unsigned num_ops, unsigned pc,
	const virg_tablet_meta *tab, void *out, size_t out_bytes, char *err, size_t errlen)
{
	if (out_bytes < sizeof(vg_prog))
		return qerr(err, errlen, "output buffer smaller than vg_prog (%zu bytes)", sizeof(vg_prog));
	return vg_decode(ops, num_ops, pc, tab, nullptr, (vg_prog *)out, err, errlen);
}

/* ------------------------------------------------------------- vg_batch */
/* Many prepared queries replayed as one HIP graph.  Each query's launch
 * pair is captured on its own stream, forked from and joined back to the
 * batch's stream, so the graph runs queries concurrently (a query over a
 * small table fills only a few compute units) and the host submits one
 * graph instead of two launches per query.  Everything a capture must not
 * contain (allocation, memset) happens before capturing: each query's
 * launch resources are sized when the batch is built. */
struct vg_batch {
	std::vector<vg_query *> qs;
	bool refs = false;          /* the queries count this batch in batch_refs */
	hipStream_t stream = nullptr;
	hipEvent_t fork = nullptr;
	std::vector<hipEvent_t> joins;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
};

extern "C" void vg_batch_free(vg_batch *b)
{
	if (!b)
		return;
	if (b->stream)
		(void)hipStreamSynchronize(b->stream);
	if (b->refs)
		refs_drop(b->qs);
	if (b->exec) (void)hipGraphExecDestroy(b->exec);
	if (b->graph) (void)hipGraphDestroy(b->graph);
	for (hipEvent_t e : b->joins)
		(void)hipEventDestroy(e);
	if (b->fork) (void)hipEventDestroy(b->fork);
	if (b->stream) (void)hipStreamDestroy(b->stream);
	delete b;
}

extern "C" vg_batch *vg_batch_create(vg_query *const *qs, unsigned n, char *err, size_t errlen)
{
	vg_batch *b = new vg_batch();
	b->qs.assign(qs, qs + n);
	for (unsigned i = 0; i < n; i++) {
		if (!qs[i] || qs[i]->d != qs[0]->d) {
			qerr(err, errlen, "batch queries must be prepared on one device");
			vg_batch_free(b);
			return nullptr;
		}
		for (unsigned j = 0; j < i; j++)
			if (qs[j] == qs[i]) {
				/* one set of launch resources per query */
				qerr(err, errlen, "query %u appears twice in the batch", i);
				vg_batch_free(b);
				return nullptr;
			}
	}
	refs_take(b->qs);
	b->refs = true;
#define BCHK(call, what) do { if (herr((call), (what))) { qerr(err, errlen, "%s failed", what); \
	vg_batch_free(b); return nullptr; } } while (0)
	if (n)
		BCHK(hipSetDevice(qs[0]->d->device), "set device");
	BCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
	BCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "event");
	for (unsigned i = 0; i < n; i++) {
		vg_query *q = qs[i];
		if (q->p->num_tiles && ensure_lr(q->lr, q->p->num_tiles, vg_kernel_grid(&q->P, q->p->num_tiles),
				q->stream) == VIRG_FAIL) {
			qerr(err, errlen, "cannot set up query %u", i);
			vg_batch_free(b);
			return nullptr;
		}
		BCHK(hipStreamSynchronize(q->stream), "query setup");
		hipEvent_t e;
		BCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
		b->joins.push_back(e);
	}
	BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture");
	int ok = hipEventRecord(b->fork, b->stream) == hipSuccess;
	for (unsigned i = 0; i < n && ok; i++) {
		vg_query *q = qs[i];
		ok = hipStreamWaitEvent(q->stream, b->fork, 0) == hipSuccess &&
			query_launch(q, q->stream) == VIRG_SUCCESS &&
			hipEventRecord(b->joins[i], q->stream) == hipSuccess &&
			hipStreamWaitEvent(b->stream, b->joins[i], 0) == hipSuccess;
	}
	const hipError_t ec = hipStreamEndCapture(b->stream, &b->graph);
	if (!ok || ec != hipSuccess) {
		(void)hipGetLastError();
		qerr(err, errlen, "graph capture failed");
		vg_batch_free(b);
		return nullptr;
	}
	BCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
#undef BCHK
	return b;
}

/* Submit the whole batch (asynchronous). */
extern "C" int vg_batch_launch(vg_batch *b)
{
	HCHK(hipGraphLaunch(b->exec, b->stream), "graph launch");
	return VIRG_SUCCESS;
}

/* Wait for the batch; rows[i] = rows selected by query i. */
extern "C" int vg_batch_sync(vg_batch *b, unsigned long long *rows)
{
	HCHK(hipStreamSynchronize(b->stream), "batch");
	for (size_t i = 0; i < b->qs.size(); i++) {
		unsigned long long n = 0;
		if (vg_query_sync(b->qs[i], &n) == VIRG_FAIL)
			return VIRG_FAIL;
		if (rows)
			rows[i] = n;
	}
	return VIRG_SUCCESS;
}

/* Mean milliseconds per batch over `iters` replays after `warmup`, timed
 * with HIP events on the batch's stream. */
extern "C" int vg_batch_time(vg_batch *b, int warmup, int iters, double *ms)
{
	hipEvent_t e0, e1;
	HCHK(hipEventCreate(&e0), "event");
	HCHK(hipEventCreate(&e1), "event");
	int ok = 1;
	for (int i = 0; i < warmup && ok; i++)
		ok = hipGraphLaunch(b->exec, b->stream) == hipSuccess;
	ok = ok && hipEventRecord(e0, b->stream) == hipSuccess;
	for (int i = 0; i < iters && ok; i++)
		ok = hipGraphLaunch(b->exec, b->stream) == hipSuccess;
	ok = ok && hipEventRecord(e1, b->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
	float t = 0;
	ok = ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess;
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (!ok) {
		herr(hipGetLastError(), "batch timing");
		return VIRG_FAIL;
	}
	*ms = iters > 0 ? t / iters : 0.0;
	return VIRG_SUCCESS;
}

/* ------------------------------------------------------------ vg_pbatch */
/* Many prepared queries run by persistent launches that work through
 * device-resident item lists.  A launch's LDS per workgroup is that of the
 * largest register file among its queries, which sets how many workgroups
 * fit on a CU; so the queries are split into groups by the occupancy their
 * own register files allow, and each group gets its own count and write
 * launches, sized to that occupancy.  The reset launch and every group's
 * pair, each group on its own branch, are captured once as a HIP graph.
 * Item size, write order, grid size and the grouping are the tunables
 * (vg_pbatch_tune). */
extern "C" size_t vg_kernel_batch_lds(int pass, const vg_prog *P);
extern "C" uint32_t vg_kernel_batch_grid(int pass, size_t lds);
extern "C" int vg_kernel_batch_reset(const vg_batch_args *A, hipStream_t stream);
extern "C" int vg_kernel_batch_passes(const vg_batch_args *A, uint32_t cgrid, size_t clds, uint32_t wgrid,
	size_t wlds, hipStream_t stream);
extern "C" int vg_kernel_batch_single(const vg_batch_args *A, uint32_t grid, size_t lds, hipStream_t stream);

/* polls of an unpublished predecessor tile before a single-pass item gives
 * up (the query then reports VG_OVF_WAIT and must be run in two passes) */
#define VG_SPIN_LIMIT (1u << 22)

/* pinned host memory that grows (by doubling) but is never shrunk */
struct pinned_buf {
	void *p = nullptr;
	size_t cap = 0;
	void *get(size_t bytes, unsigned flags = hipHostMallocDefault)
	{
		if (bytes > cap) {
			if (p)
				(void)hipHostFree(p);
			p = nullptr;
			size_t c = cap ? cap : 4096;
			while (c < bytes)
				c *= 2;
			if (hipHostMalloc(&p, c, flags) != hipSuccess) {
				(void)hipGetLastError();
				p = nullptr;
				cap = 0;
				return nullptr;
			}
			cap = c;
		}
		return p;
	}
	void release()
	{
		if (p)
			(void)hipHostFree(p);
		p = nullptr;
		cap = 0;
	}
};

/* one group of queries with its own launch pair */
struct pgroup {
	std::vector<unsigned> qids;
	size_t clds = 0, wlds = 0;
	uint32_t cgrid = 0, wgrid = 0;
	vg_batch_args A;
};

struct vg_pbatch {
	std::vector<vg_query *> qs;
	vgd_device *d = nullptr;
	hipStream_t stream = nullptr;
	std::vector<hipStream_t> branch;        /* one per group (graph capture) */
	std::vector<hipEvent_t> join;
	hipEvent_t fork = nullptr;
	/* what every run reads, uploaded with one copy per build: programs,
	 * launch arguments, per-query info, group query lists and item lists,
	 * each at its offset in d_blob (staged in h_blob) */
	char *d_blob = nullptr;
	size_t blob_cap = 0;
	pinned_buf h_blob;
	std::vector<vg_prog> progs_h;
	std::vector<vg_launch> launch_h;
	vg_prog *d_progs = nullptr;
	vg_launch *d_launch = nullptr;
	vg_bqinfo *d_info = nullptr;
	vg_bquery *d_q = nullptr;
	vg_bhead *d_head = nullptr;             /* one per group */
	size_t head_cap = 0;
	vg_item *d_items = nullptr, *d_witems = nullptr;
	uint32_t *d_icount = nullptr;
	uint64_t *d_ibase = nullptr;
	uint64_t *d_stamp = nullptr;            /* finish time of each write item */
	uint32_t *d_gq = nullptr;               /* query ids, group by group */
	size_t items_cap = 0;
	uint64_t *d_status = nullptr;           /* single pass: a word per tile */
	size_t status_cap = 0;
	std::vector<pgroup> groups;
	vg_batch_args R;                        /* the reset's arguments */
	size_t qcap = 0;                        /* queries the per-query arrays hold */
	vg_pbatch_config cfg;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	double clock_khz = 0;
	/* pinned staging: uploads and the state readback are asynchronous on
	 * the batch's stream (a synchronous copy would wait behind whatever
	 * holds the compute units) */
	/* every query's outcome, stored there by the kernels (coherent, so it
	 * can be read once the stream is idle) */
	pinned_buf h_res;
	/* config.jit: the batch's module and the device state of its launch */
	vgx_kernel *jk = nullptr;
	vgxb_args JA;
	unsigned jgrid = 0;
	vgx_launch_args *d_qa = nullptr;
	uint32_t *d_fn = nullptr;
	vgxb_item *d_jitems = nullptr;
	vgx_tile *d_jtiles = nullptr;
	uint64_t *d_jstate = nullptr;          /* counter, epoch, start; then status words */
	uint32_t *d_jdone = nullptr;
	vg_ctrl *d_jctrl = nullptr;            /* per query, the batch's own (eager runs keep theirs) */
};

static void pbatch_free_jit(vg_pbatch *b)
{
	void *bufs[] = { b->d_qa, b->d_fn, b->d_jitems, b->d_jtiles, b->d_jstate, b->d_jdone, b->d_jctrl };
	for (void *p : bufs)
		if (p)
			(void)hipFree(p);
	b->d_qa = nullptr;
	b->d_fn = nullptr;
	b->d_jitems = nullptr;
	b->d_jtiles = nullptr;
	b->d_jstate = nullptr;
	b->d_jdone = nullptr;
	b->d_jctrl = nullptr;
	b->jk = nullptr;
}

static void pbatch_drop_graph(vg_pbatch *b)
{
	if (b->exec) (void)hipGraphExecDestroy(b->exec);
	if (b->graph) (void)hipGraphDestroy(b->graph);
	b->exec = nullptr;
	b->graph = nullptr;
}

extern "C" void vg_pbatch_free(vg_pbatch *b)
{
	if (!b)
		return;
	if (b->stream)
		(void)hipStreamSynchronize(b->stream);
	refs_drop(b->qs);
	pbatch_drop_graph(b);
	pbatch_free_jit(b);
	void *bufs[] = { b->d_blob, b->d_q, b->d_head, b->d_icount, b->d_ibase, b->d_status, b->d_stamp };
	for (void *p : bufs)
		if (p)
			(void)hipFree(p);
	for (hipEvent_t e : b->join)
		(void)hipEventDestroy(e);
	for (hipStream_t s : b->branch)
		(void)hipStreamDestroy(s);
	pinned_buf *pins[] = { &b->h_blob, &b->h_res };
	for (pinned_buf *pb : pins)
		pb->release();
	if (b->fork)
		(void)hipEventDestroy(b->fork);
	if (b->stream)
		(void)hipStreamDestroy(b->stream);
	delete b;
}

/* the kernel arguments of one query's launch (as launch() builds them) */
static void fill_launch(const vg_query *q, vg_launch &L)
{
	memset(&L, 0, sizeof L);
	L.tabs = q->p->d_desc;
	L.tile_tab = q->p->d_tile;
	L.num_tiles = q->p->num_tiles;
	L.capacity = q->capacity;
	L.ctrl = q->lr.ctrl;
	L.vbits = q->lr.vbits;
	L.tcount = q->lr.tcount;
	L.wgcount = q->lr.wgcount;
	for (unsigned j = 0; j < q->P.nout; j++)
		L.out[j] = q->out[j];
}

/* Issue a run on b->stream: the reset, then every group's launches, each
 * group on its own branch when there are several. */
static int pbatch_record(vg_pbatch *b)
{
	if (b->jk)
		return vgxb_launch(b->jk, &b->JA, b->jgrid, b->stream) == 0;
	const size_t ng = b->groups.size();
	const bool single = b->cfg.single != 0;
	int rc = vg_kernel_batch_reset(&b->R, b->stream) == 0 && hipEventRecord(b->fork, b->stream) == hipSuccess;
	for (size_t gi = 0; gi < ng && rc; gi++) {
		pgroup &g = b->groups[gi];
		if (g.A.nitems == 0)
			continue;
		hipStream_t s = ng == 1 ? b->stream : b->branch[gi];
		rc = (ng == 1 || hipStreamWaitEvent(s, b->fork, 0) == hipSuccess) &&
			(single ? vg_kernel_batch_single(&g.A, g.wgrid, g.wlds, s)
			        : vg_kernel_batch_passes(&g.A, g.cgrid, g.clds, g.wgrid, g.wlds, s)) == 0 &&
			(ng == 1 || (hipEventRecord(b->join[gi], s) == hipSuccess &&
				hipStreamWaitEvent(b->stream, b->join[gi], 0) == hipSuccess));
	}
	return rc;
}

/* Capture the batch's launches as its graph (not for direct batches). */
static int pbatch_capture(vg_pbatch *b, char *err, size_t errlen)
{
	pbatch_drop_graph(b);
	if (b->cfg.direct)
		return VIRG_SUCCESS;
	if (herr(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture"))
		return qerr(err, errlen, "graph capture failed");
	const int rc = pbatch_record(b);
	const hipError_t ec = hipStreamEndCapture(b->stream, &b->graph);
	if (!rc || ec != hipSuccess) {
		(void)hipGetLastError();
		return qerr(err, errlen, "graph capture failed");
	}
	if (herr(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate"))
		return qerr(err, errlen, "graph instantiate failed");
	return VIRG_SUCCESS;
}

/* config.jit: compile the batch's programs into one module, lay out each
 * query's launch (tile table, status words) and the work items (every
 * query's chunks in order, the smallest query first: the many small
 * queries of a mixed batch complete early, the large ones spread over the
 * whole device at the end), upload them. */
static int pbatch_build_jit(vg_pbatch *b, char *err, size_t errlen)
{
	pbatch_free_jit(b);
	const unsigned n = (unsigned)b->qs.size();
	if (n == 0)
		return pbatch_capture(b, err, errlen);
	std::vector<const vg_prog *> progs(n);
	std::vector<int> aligned(n);
	for (unsigned i = 0; i < n; i++) {
		progs[i] = &b->qs[i]->P;
		bool al = true;
		for (const vg_tabdesc &d : b->qs[i]->p->desc)
			al = al && d.aligned;
		aligned[i] = al ? 1 : 0;
	}
	std::vector<uint32_t> fn(n);
	vgx_kernel *k = vgxb_get(progs.data(), aligned.data(), n, fn.data(), err, errlen);
	if (!k)
		return VIRG_FAIL;
	std::vector<vgx_launch_args> qa(n);
	std::vector<vgx_tile> tiles;
	std::vector<uint32_t> nchunks(n);
	size_t status_words = 0;
	std::vector<size_t> toff(n), soff(n);
	for (unsigned i = 0; i < n; i++) {
		vg_query *q = b->qs[i];
		vgx_shape sh;
		if (vgx_shape_of(&q->P, aligned[i], 0, 0, vgxb_shape_mode(), &sh, err, errlen) != 0)
			return VIRG_FAIL;
		const std::vector<vgx_tile> t = jit_tiles(*q->p, sh.TR);
		toff[i] = tiles.size();
		tiles.insert(tiles.end(), t.begin(), t.end());
		nchunks[i] = (uint32_t)((t.size() + sh.S - 1) / sh.S);
		soff[i] = status_words;
		status_words += nchunks[i];
	}
	std::vector<unsigned> order(n);
	for (unsigned i = 0; i < n; i++)
		order[i] = i;
	std::stable_sort(order.begin(), order.end(), [&](unsigned x, unsigned y) { return nchunks[x] < nchunks[y]; });
	std::vector<vgxb_item> items;
	for (unsigned i : order)
		for (uint32_t c = 0; c < nchunks[i]; c++)
			items.push_back(vgxb_item{i, c});
	const size_t state_bytes = (3 + status_words) * sizeof(uint64_t);
	if (herr(hipMalloc((void **)&b->d_qa, n * sizeof(vgx_launch_args)), "batch") ||
		herr(hipMalloc((void **)&b->d_fn, n * sizeof(uint32_t)), "batch") ||
		herr(hipMalloc((void **)&b->d_jitems, (items.size() ? items.size() : 1) * sizeof(vgxb_item)), "batch") ||
		herr(hipMalloc((void **)&b->d_jtiles, (tiles.size() ? tiles.size() : 1) * sizeof(vgx_tile)), "batch") ||
		herr(hipMalloc((void **)&b->d_jstate, state_bytes), "batch") ||
		herr(hipMalloc((void **)&b->d_jdone, n * sizeof(uint32_t)), "batch") ||
		herr(hipMalloc((void **)&b->d_jctrl, n * sizeof(vg_ctrl)), "batch"))
		return qerr(err, errlen, "cannot allocate the batch");
	for (unsigned i = 0; i < n; i++) {
		vg_query *q = b->qs[i];
		vgx_launch_args &a = qa[i];
		memset(&a, 0, sizeof a);
		a.tabs = q->p->d_desc;
		a.tiles = b->d_jtiles + toff[i];
		a.num_tiles = (uint32_t)(i + 1 < n ? toff[i + 1] - toff[i] : tiles.size() - toff[i]);
		a.nchunks = nchunks[i];
		a.status = b->d_jstate + 3 + soff[i];
		a.ctrl = b->d_jctrl + i;
		a.capacity = q->capacity;
		a.spin_limit = 1u << 20;
		for (unsigned j = 0; j < q->P.nout; j++)
			a.out[j] = q->out[j];
		vgx_immediates(&q->P, &a);
	}
	if (herr(hipMemcpy(b->d_qa, qa.data(), n * sizeof(vgx_launch_args), hipMemcpyHostToDevice), "batch") ||
		herr(hipMemcpy(b->d_fn, fn.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice), "batch") ||
		(items.size() && herr(hipMemcpy(b->d_jitems, items.data(), items.size() * sizeof(vgxb_item),
			hipMemcpyHostToDevice), "batch")) ||
		(tiles.size() && herr(hipMemcpy(b->d_jtiles, tiles.data(), tiles.size() * sizeof(vgx_tile),
			hipMemcpyHostToDevice), "batch")) ||
		herr(hipMemset(b->d_jstate, 0, state_bytes), "batch") ||
		herr(hipMemset(b->d_jctrl, 0, n * sizeof(vg_ctrl)), "batch"))
		return qerr(err, errlen, "cannot upload the batch");
	if (b->h_res.p)
		memset(b->h_res.p, 0, n * sizeof(vg_bresult));
	vg_bresult *d_res = nullptr;
	if (b->h_res.p && herr(hipHostGetDevicePointer((void **)&d_res, b->h_res.p, 0), "batch results"))
		return qerr(err, errlen, "cannot map the batch's results");
	vgxb_args &A = b->JA;
	memset(&A, 0, sizeof A);
	A.qa = b->d_qa;
	A.fn = b->d_fn;
	A.items = b->d_jitems;
	A.nitems = (uint32_t)items.size();
	A.nquery = n;
	A.state = b->d_jstate;
	A.done = b->d_jdone;
	A.hres = d_res;
	b->jgrid = vgxb_grid(k, A.nitems, (unsigned)b->d->prop.multiProcessorCount);
	b->jk = k;
	b->groups.clear();
	return pbatch_capture(b, err, errlen);
}

/* Group the queries (b->cfg.split 0: by the write launch's occupancy,
 * 1: all together), build the item lists, upload them and capture the
 * graph. */
static int pbatch_build(vg_pbatch *b, char *err, size_t errlen)
{
	if (b->cfg.jit && !b->cfg.single)
		return pbatch_build_jit(b, err, errlen);
	pbatch_free_jit(b);
	const unsigned n = (unsigned)b->qs.size();
	const uint32_t tpi = b->cfg.tiles_per_item ? b->cfg.tiles_per_item : 8;
	/* groups, keyed by the write launch's workgroups per device */
	std::vector<uint32_t> key(n, 0);
	if (b->cfg.split == 0)
		for (unsigned i = 0; i < n; i++)
			key[i] = vg_kernel_batch_grid(1, vg_kernel_batch_lds(1, &b->qs[i]->P));
	std::vector<uint32_t> keys(key);
	std::sort(keys.begin(), keys.end());
	keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
	std::reverse(keys.begin(), keys.end());              /* highest occupancy first */
	b->groups.assign(keys.size(), pgroup());
	for (unsigned i = 0; i < n; i++) {
		const size_t g = std::find(keys.begin(), keys.end(), key[i]) - keys.begin();
		b->groups[g].qids.push_back(i);
	}
	std::vector<vg_bqinfo> info(n);
	uint32_t ioff = 0, soff = 0;
	for (unsigned i = 0; i < n; i++) {
		const uint32_t nt = b->qs[i]->p->num_tiles;
		info[i].nitem = (nt + tpi - 1) / tpi;
		info[i].ioff = ioff;
		info[i].soff = soff;           /* single pass: a status word per item */
		info[i].wstart = 0;
		ioff += info[i].nitem;
		soff += info[i].nitem;
	}
	const bool single = b->cfg.single != 0;
	if (single && soff > b->status_cap) {
		const size_t want = soff > 2 * b->status_cap ? soff : 2 * b->status_cap;
		if (b->d_status) (void)hipFree(b->d_status);
		b->d_status = nullptr;
		b->status_cap = 0;
		if (herr(hipMalloc((void **)&b->d_status, want * sizeof(uint64_t)), "tile status"))
			return qerr(err, errlen, "cannot allocate the batch's tile status words");
		b->status_cap = want;
	}
	auto tiles = [&](unsigned q) { return b->qs[q]->p->num_tiles; };
	std::vector<vg_item> items, witems;
	auto add = [&](std::vector<vg_item> &v, unsigned q) {
		const uint32_t nt = tiles(q);
		for (uint32_t k = 0; k < info[q].nitem; k++)
			v.push_back({q, k * tpi, (k + 1) * tpi < nt ? (k + 1) * tpi : nt, k});
	};
	std::vector<size_t> goff;
	for (pgroup &g : b->groups) {
		/* count list: largest query first (small items last even out the
		 * end); write list in the configured order */
		std::vector<unsigned> by_size(g.qids), worder(g.qids);
		std::stable_sort(by_size.begin(), by_size.end(), [&](unsigned x, unsigned y) {
			return tiles(x) > tiles(y); });
		if (b->cfg.order == 1)
			worder.assign(by_size.rbegin(), by_size.rend());
		else if (b->cfg.order == 2)
			worder = by_size;
		goff.push_back(items.size());
		for (unsigned q : by_size)
			add(items, q);
		for (unsigned q : worder) {
			/* within the group's list: stamps are indexed the same way */
			info[q].wstart = (uint32_t)(witems.size() - goff.back());
			add(witems, q);
		}
		for (unsigned q : g.qids) {
			const size_t c = vg_kernel_batch_lds(0, &b->qs[q]->P), w = vg_kernel_batch_lds(1, &b->qs[q]->P);
			g.clds = c > g.clds ? c : g.clds;
			g.wlds = w > g.wlds ? w : g.wlds;
		}
		if (single)
			g.clds = 0;
	}
	goff.push_back(items.size());
	const size_t ni = items.size();
	if (ni > b->items_cap) {
		const size_t want = ni > 2 * b->items_cap ? ni : 2 * b->items_cap;
		if (b->d_icount) (void)hipFree(b->d_icount);
		if (b->d_ibase) (void)hipFree(b->d_ibase);
		if (b->d_stamp) (void)hipFree(b->d_stamp);
		b->d_icount = nullptr;
		b->d_ibase = nullptr;
		b->d_stamp = nullptr;
		b->items_cap = 0;
		if (herr(hipMalloc((void **)&b->d_icount, want * sizeof(uint32_t)), "batch counts") ||
			herr(hipMalloc((void **)&b->d_ibase, want * sizeof(uint64_t)), "batch offsets") ||
			herr(hipMalloc((void **)&b->d_stamp, want * sizeof(uint64_t)), "batch stamps"))
			return qerr(err, errlen, "cannot allocate the batch's item lists");
		b->items_cap = want;
	}
	const size_t ng = b->groups.size();
	if (ng > b->head_cap) {
		if (b->d_head) (void)hipFree(b->d_head);
		b->d_head = nullptr;
		b->head_cap = 0;
		if (herr(hipMalloc((void **)&b->d_head, ng * sizeof(vg_bhead)), "batch heads"))
			return qerr(err, errlen, "cannot allocate the batch's list heads");
		b->head_cap = ng;
	}
	while (b->branch.size() < ng) {
		hipStream_t s;
		hipEvent_t e;
		if (herr(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream"))
			return qerr(err, errlen, "cannot create the batch's streams");
		b->branch.push_back(s);
		if (herr(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event"))
			return qerr(err, errlen, "cannot create the batch's events");
		b->join.push_back(e);
	}
	{
		/* one upload: programs, launch arguments, info, group lists, items */
		auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
		const size_t o_launch = al(n * sizeof(vg_prog));
		const size_t o_info = o_launch + al(n * sizeof(vg_launch));
		const size_t o_gq = o_info + al(n * sizeof(vg_bqinfo));
		const size_t o_items = o_gq + al(n * sizeof(uint32_t));
		const size_t o_witems = o_items + (single ? 0 : al(ni * sizeof(vg_item)));
		const size_t total = o_witems + al(ni * sizeof(vg_item));
		if (total > b->blob_cap) {
			const size_t want = total > 2 * b->blob_cap ? total : 2 * b->blob_cap;
			if (b->d_blob) (void)hipFree(b->d_blob);
			b->d_blob = nullptr;
			b->blob_cap = 0;
			if (herr(hipMalloc((void **)&b->d_blob, want), "batch"))
				return qerr(err, errlen, "cannot allocate the batch");
			b->blob_cap = want;
		}
		char *h = (char *)b->h_blob.get(total);
		if (!h)
			return qerr(err, errlen, "cannot allocate pinned staging");
		memcpy(h, b->progs_h.data(), n * sizeof(vg_prog));
		memcpy(h + o_launch, b->launch_h.data(), n * sizeof(vg_launch));
		memcpy(h + o_info, info.data(), n * sizeof(vg_bqinfo));
		uint32_t *hg = (uint32_t *)(h + o_gq);
		size_t k = 0;
		for (pgroup &g : b->groups)
			for (unsigned q : g.qids)
				hg[k++] = q;
		if (!single && ni)
			memcpy(h + o_items, items.data(), ni * sizeof(vg_item));
		if (ni)
			memcpy(h + o_witems, witems.data(), ni * sizeof(vg_item));
		b->d_progs = (vg_prog *)b->d_blob;
		b->d_launch = (vg_launch *)(b->d_blob + o_launch);
		b->d_info = (vg_bqinfo *)(b->d_blob + o_info);
		b->d_gq = (uint32_t *)(b->d_blob + o_gq);
		b->d_items = (vg_item *)(b->d_blob + (single ? o_witems : o_items));
		b->d_witems = (vg_item *)(b->d_blob + o_witems);
		if (herr(hipMemcpyAsync(b->d_blob, h, total, hipMemcpyHostToDevice, b->stream), "batch"))
			return qerr(err, errlen, "cannot upload the batch");
		/* graph replays read the lists later: they must have landed */
		if (!b->cfg.direct && herr(hipStreamSynchronize(b->stream), "batch items"))
			return qerr(err, errlen, "cannot upload the batch's item lists");
	}
	/* the kernels store each query's outcome through the device's view of
	 * the pinned results */
	vg_bresult *d_res = nullptr;
	if (b->h_res.p && herr(hipHostGetDevicePointer((void **)&d_res, b->h_res.p, 0), "batch results"))
		return qerr(err, errlen, "cannot map the batch's results");
	size_t gq_off = 0;
	for (size_t gi = 0; gi < ng; gi++) {
		pgroup &g = b->groups[gi];
		vg_batch_args &A = g.A;
		memset(&A, 0, sizeof A);
		A.progs = b->d_progs;
		A.launch = b->d_launch;
		A.info = b->d_info;
		A.items = b->d_items + goff[gi];
		A.witems = b->d_witems + goff[gi];
		A.nitems = (uint32_t)(goff[gi + 1] - goff[gi]);
		A.nquery = n;
		A.q = b->d_q;
		A.icount = b->d_icount;
		A.ibase = b->d_ibase;
		A.stamp = b->d_stamp + goff[gi];
		A.gq = b->d_gq + gq_off;
		A.ngq = (uint32_t)g.qids.size();
		gq_off += g.qids.size();
		A.head = b->d_head + gi;
		A.nhead = 1;
		A.spin_limit = VG_SPIN_LIMIT;
		A.hres = d_res;
		A.status = single ? b->d_status : nullptr;
		A.nstatus = 0;
		const uint32_t cmax = vg_kernel_batch_grid(0, g.clds), wmax = vg_kernel_batch_grid(single ? 2 : 1, g.wlds);
		const unsigned fr = b->cfg.grid_div ? b->cfg.grid_div : 1;
		g.cgrid = cmax / fr ? cmax / fr : 1;
		g.wgrid = wmax / fr ? wmax / fr : 1;
		const uint32_t most = A.nitems ? A.nitems : 1;
		g.cgrid = g.cgrid < most ? g.cgrid : most;
		g.wgrid = g.wgrid < most ? g.wgrid : most;
	}
	/* the reset clears every query and every group's head */
	b->R = b->groups[0].A;
	b->R.head = b->d_head;
	b->R.nhead = (uint32_t)ng;
	b->R.status = single ? b->d_status : nullptr;
	b->R.nstatus = single ? soff : 0;
	return pbatch_capture(b, err, errlen);
}

/* Take a new set of queries: check them, make sure each has what the
 * configured mode needs, and upload their programs and launch arguments. */
static int pbatch_set_queries(vg_pbatch *b, vg_query *const *qs, unsigned n, char *err, size_t errlen)
{
	for (unsigned i = 0; i < n; i++) {
		if (!qs[i] || qs[i]->d != b->d)
			return qerr(err, errlen, "batch queries must be prepared on one device");
		for (unsigned j = 0; j < i; j++)
			if (qs[j] == qs[i])
				return qerr(err, errlen, "query %u appears twice in the batch", i);
	}
	refs_drop(b->qs);
	b->qs.assign(qs, qs + n);
	refs_take(b->qs);
	std::vector<vg_prog> progs(n);
	std::vector<vg_launch> launches(n);
	for (unsigned i = 0; i < n; i++) {
		vg_query *q = qs[i];
		const uint32_t nt = q->p->num_tiles;
		/* both forms keep validity bytes and strip counts per tile */
		if (q->pooled_ctrl)
			take_pooled_lr(q->d, q->lr, nt, vg_kernel_grid(&q->P, nt ? nt : 1));
		if (ensure_lr(q->lr, nt, vg_kernel_grid(&q->P, nt ? nt : 1), q->stream) == VIRG_FAIL)
			return qerr(err, errlen, "cannot set up query %u", i);
		progs[i] = q->P;
		fill_launch(q, launches[i]);
	}
	for (unsigned i = 0; i < n; i++)
		if (!qs[i]->pooled_ctrl && herr(hipStreamSynchronize(qs[i]->stream), "query setup"))
			return qerr(err, errlen, "query setup failed");
	if (n > b->qcap) {
		const size_t cap = n < 64 ? 64 : n > 2 * b->qcap ? n : 2 * b->qcap;
		if (b->d_q)
			(void)hipFree(b->d_q);
		b->d_q = nullptr;
		b->qcap = 0;
		if (herr(hipMalloc((void **)&b->d_q, cap * sizeof(vg_bquery)), "batch"))
			return qerr(err, errlen, "cannot allocate the batch");
		b->qcap = cap;
	}
	if (n && !b->h_res.get(n * sizeof(vg_bresult), hipHostMallocCoherent))
		return qerr(err, errlen, "cannot allocate pinned staging");
	/* uploaded with the item lists when the batch is built */
	b->progs_h.swap(progs);
	b->launch_h.swap(launches);
	return VIRG_SUCCESS;
}

extern "C" vg_pbatch *vg_pbatch_create(vg_query *const *qs, unsigned n, const vg_pbatch_config *cfg,
	char *err, size_t errlen)
{
	vg_pbatch_config c0;
	memset(&c0, 0, sizeof c0);
	if (cfg)
		c0 = *cfg;
	if (n == 0 && !c0.direct) {
		qerr(err, errlen, "empty batch");
		return nullptr;
	}
	vg_pbatch *b = new vg_pbatch();
	b->cfg = c0;
#define PCHK(cond, ...) do { if (cond) { (void)hipGetLastError(); qerr(err, errlen, __VA_ARGS__); \
	vg_pbatch_free(b); return nullptr; } } while (0)
	if (n) {
		b->d = qs[0] ? qs[0]->d : nullptr;
		PCHK(!b->d, "batch queries must be prepared on one device");
	} else {
		b->d = dev_of(c0.db);
		PCHK(!b->d, "no usable GPU");
	}
	PCHK(hipSetDevice(b->d->device) != hipSuccess, "set device");
	PCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess, "stream");
	PCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess, "event");
	int khz = 0;
	PCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->d->device) != hipSuccess || khz <= 0,
		"device clock rate");
	b->clock_khz = khz;
#undef PCHK
	if (pbatch_set_queries(b, qs, n, err, errlen) == VIRG_FAIL ||
		(n && pbatch_build(b, err, errlen) == VIRG_FAIL)) {
		vg_pbatch_free(b);
		return nullptr;
	}
	return b;
}

/* Replace the batch's queries (waits for the batch's last run first). */
extern "C" int vg_pbatch_reload(vg_pbatch *b, vg_query *const *qs, unsigned n, char *err, size_t errlen)
{
	HCHK(hipStreamSynchronize(b->stream), "batch");
	if (pbatch_set_queries(b, qs, n, err, errlen) == VIRG_FAIL)
		return VIRG_FAIL;
	if (n == 0) {
		b->groups.clear();
		pbatch_drop_graph(b);
		return VIRG_SUCCESS;
	}
	return pbatch_build(b, err, errlen);
}

/* The stream a batch runs on, and where its kernels store each query's
 * outcome (pinned host memory, `done` last), for the serving loop. */
extern "C" hipStream_t vgd_pbatch_stream(vg_pbatch *b)
{
	return b->stream;
}

extern "C" const vg_bresult *vgd_pbatch_outcomes(vg_pbatch *b)
{
	return (const vg_bresult *)b->h_res.p;
}

/* 1 when the batch's last run has finished (does not wait). */
extern "C" int vg_pbatch_done(vg_pbatch *b)
{
	const hipError_t e = hipStreamQuery(b->stream);
	if (e == hipSuccess)
		return 1;
	if (e != hipErrorNotReady)
		(void)hipGetLastError();
	return 0;
}

extern "C" int vg_pbatch_configure(vg_pbatch *b, const vg_pbatch_config *cfg, char *err, size_t errlen)
{
	HCHK(hipStreamSynchronize(b->stream), "batch");
	b->cfg = *cfg;
	return pbatch_build(b, err, errlen);
}

extern "C" int vg_pbatch_get_config(vg_pbatch *b, vg_pbatch_config *cfg, unsigned *ngroups)
{
	*cfg = b->cfg;
	if (ngroups)
		*ngroups = (unsigned)b->groups.size();
	return VIRG_SUCCESS;
}

extern "C" int vg_pbatch_launch(vg_pbatch *b)
{
	if (b->qs.empty())
		return VIRG_SUCCESS;
	if (b->cfg.direct) {
		if (!pbatch_record(b)) {
			herr(hipGetLastError(), "batch launch");
			return VIRG_FAIL;
		}
		return VIRG_SUCCESS;
	}
	HCHK(hipGraphLaunch(b->exec, b->stream), "graph launch");
	return VIRG_SUCCESS;
}

/* Copy the batch's per-query state back (after its run) and read it. */
static int pbatch_state(vg_pbatch *b, unsigned long long *rows, double *latency_ms, unsigned *flags)
{
	const size_t n = b->qs.size();
	/* idle after a run that finished: nothing to copy, the kernels stored
	 * every query's outcome in h_res */
	HCHK(hipStreamSynchronize(b->stream), "batch");
	const vg_bresult *hr = (const vg_bresult *)b->h_res.p;
	if (n && !hr)
		return VIRG_FAIL;
	for (size_t i = 0; i < n; i++) {
		const vg_bresult r = hr[i];
		if (latency_ms)
			latency_ms[i] = r.t_done > r.t_start ? (double)(r.t_done - r.t_start) / b->clock_khz : 0.0;
		if (rows)
			rows[i] = r.total;
		if (flags)
			flags[i] = r.flags;
		b->qs[i]->last_rows = r.flags ? 0 : r.total;
	}
	return VIRG_SUCCESS;
}

/* Wait for the run; rows[i] = rows query i selected, latency_ms[i] = its
 * results complete, from the start of the run (device clock); either may
 * be NULL.  Fails when a query's result exceeded its capacity. */
extern "C" int vg_pbatch_sync(vg_pbatch *b, unsigned long long *rows, double *latency_ms)
{
	const size_t n = b->qs.size();
	std::vector<unsigned> flags(n);
	if (n == 0)
		return VIRG_SUCCESS;
	if (pbatch_state(b, rows, latency_ms, flags.data()) == VIRG_FAIL)
		return VIRG_FAIL;
	for (size_t i = 0; i < n; i++)
		VIRG_CHECK(flags[i] != 0, "result exceeded the output capacity");
	return VIRG_SUCCESS;
}

/* Wait for the run; per query: rows selected (valid when flags[i] is 0),
 * latency, and flags[i] = the VG_OVF_* bits (the result is then
 * incomplete and the query must be run again with exact sizing). */
extern "C" int vg_pbatch_results(vg_pbatch *b, unsigned long long *rows, double *latency_ms, unsigned *flags)
{
	if (b->qs.empty())
		return VIRG_SUCCESS;
	return pbatch_state(b, rows, latency_ms, flags);
}

/* mean ms per run of `iters` graph replays after `warmup`, HIP events */
static int graph_time(hipGraphExec_t exec, hipStream_t s, int warmup, int iters, double *ms)
{
	hipEvent_t e0, e1;
	HCHK(hipEventCreate(&e0), "event");
	if (herr(hipEventCreate(&e1), "event")) {
		(void)hipEventDestroy(e0);
		return VIRG_FAIL;
	}
	int ok = 1;
	for (int i = 0; i < warmup && ok; i++)
		ok = hipGraphLaunch(exec, s) == hipSuccess;
	ok = ok && hipEventRecord(e0, s) == hipSuccess;
	for (int i = 0; i < iters && ok; i++)
		ok = hipGraphLaunch(exec, s) == hipSuccess;
	ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
	float t = 0;
	ok = ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess;
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (!ok) {
		herr(hipGetLastError(), "graph timing");
		return VIRG_FAIL;
	}
	*ms = iters > 0 ? t / iters : 0.0;
	return VIRG_SUCCESS;
}

extern "C" int vg_pbatch_time(vg_pbatch *b, int warmup, int iters, double *ms)
{
	if (b->cfg.direct || b->qs.empty())
		return VIRG_FAIL;
	return graph_time(b->exec, b->stream, warmup, iters, ms);
}

/* Autotuner: time every candidate configuration (tiles per item x write
 * order x grid size) over `iters` replays on this batch and keep the
 * fastest; *best (optional) receives it, *best_ms its time. */
extern "C" int vg_pbatch_tune(vg_pbatch *b, int iters, vg_pbatch_config *best, double *best_ms,
	char *err, size_t errlen)
{
	static const unsigned tpis[] = { 1, 2, 4, 8, 16, 32, 64 };
	static const unsigned orders[] = { 0, 1, 2 };
	static const unsigned divs[] = { 1, 2 };
	static const unsigned splits[] = { 0, 1 };
	vg_pbatch_config top;
	memset(&top, 0, sizeof top);
	double top_ms = 1e30;
	for (unsigned sp : splits)
		for (unsigned t : tpis)
			for (unsigned o : orders)
				for (unsigned f : divs) {
					vg_pbatch_config c;
					memset(&c, 0, sizeof c);
					c.tiles_per_item = t;
					c.order = o;
					c.grid_div = f;
					c.split = sp;
					double ms = 0;
					if (vg_pbatch_configure(b, &c, err, errlen) == VIRG_FAIL ||
						vg_pbatch_time(b, 1, iters, &ms) == VIRG_FAIL)
						return VIRG_FAIL;
					if (ms < top_ms) {
						top_ms = ms;
						top = c;
					}
				}
	if (vg_pbatch_configure(b, &top, err, errlen) == VIRG_FAIL)
		return VIRG_FAIL;
	if (best)
		*best = top;
	if (best_ms)
		*best_ms = top_ms;
	return VIRG_SUCCESS;
}

/* The plain way: each query's count and write launches, one query after
 * another on one stream.  Mean ms per round over `iters` rounds after
 * `warmup`; latency_ms[i] (optional) = query i's results complete, from
 * the start of the last round (HIP events). */
extern "C" int vg_plain_time(vg_query *const *qs, unsigned n, int warmup, int iters, double *ms,
	double *latency_ms)
{
	if (n == 0)
		return VIRG_FAIL;
	HCHK(hipSetDevice(qs[0]->d->device), "set device");
	hipStream_t s;
	HCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
	std::vector<hipEvent_t> ev(n + 2, nullptr);
	int ok = 1;
	for (auto &e : ev)
		ok = ok && hipEventCreate(&e) == hipSuccess;
	for (unsigned i = 0; i < n && ok; i++)
		ok = ensure_lr(qs[i]->lr, qs[i]->p->num_tiles, vg_kernel_grid(&qs[i]->P, qs[i]->p->num_tiles ?
			qs[i]->p->num_tiles : 1), qs[i]->stream) == VIRG_SUCCESS &&
			hipStreamSynchronize(qs[i]->stream) == hipSuccess;
	auto round = [&](bool mark) {
		for (unsigned i = 0; i < n && ok; i++) {
			ok = query_launch(qs[i], s) == VIRG_SUCCESS;
			if (mark)
				ok = ok && hipEventRecord(ev[i + 2], s) == hipSuccess;
		}
	};
	for (int w = 0; w < warmup && ok; w++)
		round(false);
	ok = ok && hipEventRecord(ev[0], s) == hipSuccess;
	for (int it = 0; it + 1 < iters && ok; it++)
		round(false);
	ok = ok && hipEventRecord(ev[1], s) == hipSuccess;   /* start of the last round */
	round(true);
	ok = ok && hipStreamSynchronize(s) == hipSuccess;
	float t = 0;
	ok = ok && hipEventElapsedTime(&t, ev[0], ev[n + 1]) == hipSuccess;
	if (ok) {
		*ms = iters > 0 ? t / iters : 0.0;
		for (unsigned i = 0; i < n && ok && latency_ms; i++) {
			float l = 0;
			ok = hipEventElapsedTime(&l, ev[1], ev[i + 2]) == hipSuccess;
			latency_ms[i] = l;
		}
	}
	for (auto &e : ev)
		if (e)
			(void)hipEventDestroy(e);
	(void)hipStreamDestroy(s);
	if (!ok) {
		herr(hipGetLastError(), "plain launches");
		return VIRG_FAIL;
	}
	return VIRG_SUCCESS;
}

/* Results of `q` go to caller device memory from now on: out[j] holds
 * room for `capacity` rows of result column j (its width in bytes). */
extern "C" int vg_query_set_outputs(vg_query *q, void *const *out, unsigned long long capacity)
{
	if (q->batch_refs) {
		fprintf(stderr, "::: Virginian: vg_query_set_outputs refused: %u batch(es) hold the query's "
			"launch arguments (free or reload them first)\n", q->batch_refs);
		return VIRG_FAIL;
	}
	if (q->own_out)
		for (unsigned j = 0; j < VG_MAX_OUT; j++)
			if (q->out[j]) {
				(void)hipStreamSynchronize(q->stream);
				(void)hipFree(q->out[j]);
			}
	q->own_out = false;
	for (unsigned j = 0; j < q->P.nout; j++) {
		q->out[j] = out[j];
		q->ja.out[j] = out[j];
	}
	q->capacity = capacity;
	q->ja.capacity = capacity;
	return VIRG_SUCCESS;
}

/* Bytes per row of each result column, in order; returns their count. */
extern "C" unsigned vg_query_out_widths(vg_query *q, unsigned *widths)
{
	for (unsigned j = 0; j < q->P.nout; j++)
		widths[j] = (unsigned)q->P.out[j].width;
	return q->P.nout;
}

/* Device memory for caller-managed outputs (e.g. a serving arena). */
extern "C" void *vg_device_alloc(virginian *v, size_t bytes)
{
	vgd_device *d = dev_of(v);
	void *p = nullptr;
	if (!d || hipSetDevice(d->device) != hipSuccess || hipMalloc(&p, bytes) != hipSuccess) {
		(void)hipGetLastError();
		return nullptr;
	}
	return p;
}

extern "C" void vg_device_free(virginian *v, void *p)
{
	(void)v;
	if (p)
		(void)hipFree(p);
}

/* Rows the last (two-pass) launch of q selected per tablet, from the count
 * pass's per-tile counts. */
extern "C" unsigned vg_query_tablet_rows(vg_query *q, unsigned long long *rows, unsigned n)
{
	const plan &p = *q->p;
	const size_t nt = p.num_tiles;
	if (!q->lr.tcount || nt == 0) {
		for (unsigned i = 0; i < n && i < p.desc.size(); i++)
			rows[i] = 0;
		return (unsigned)p.desc.size();
	}
	std::vector<uint32_t> tc(nt * 4);
	if (hipStreamSynchronize(q->stream) != hipSuccess ||
		hipMemcpy(tc.data(), q->lr.tcount, nt * 4 * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) {
		(void)hipGetLastError();
		return 0;
	}
	const uint32_t T = vg_kernel_tile_rows();
	for (size_t i = 0; i < p.desc.size() && i < n; i++) {
		const uint32_t f = p.desc[i].first_tile, k = (p.desc[i].rows + T - 1) / T;
		unsigned long long s = 0;
		for (uint32_t t = f; t < f + k; t++)
			s += (unsigned long long)tc[t * 4] + tc[t * 4 + 1] + tc[t * 4 + 2] + tc[t * 4 + 3];
		rows[i] = s;
	}
	return (unsigned)p.desc.size();
}

/* the launch's one-tablet descriptor and its all-zero tile table, written
 * into the scratch by a kernel (no memset or copy nodes, so the call can be
 * captured in a graph) */
__global__ __launch_bounds__(256) void vg_noalloc_setup(vg_tabdesc *desc, uint32_t *tiles, uint32_t ntiles,
	const vg_tabdesc d)
{
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if (i == 0)
		*desc = d;
	if (i < ntiles)
		tiles[i] = 0;
}

/* reference: none (a variant of vg_query_run for callers that cannot
 * allocate).  Runs the data-parallel section of ops over `rows` rows whose
 * columns are already in device memory (cols[c]: column c's row 0; key:
 * the key column, needed only by Rowid) into the caller's outputs, with
 * the interpreter's count and write passes, on `stream`, asynchronously.
 * result (device, 32 bytes, a vg_ctrl): [0] rows selected, [1] low word
 * nonzero when they exceeded `capacity`.  Scratch comes from the library's
 * per-GPU array; calls that overlap on one GPU race on it. */
extern "C" int vg_select_noalloc(const virg_op *ops, unsigned num_ops, const void *const *cols,
	const int *col_types, unsigned ncols, const void *key, int key_type, unsigned long long rows,
	void *const *out, unsigned long long capacity, void *result, void *stream_, char *err, size_t errlen)
{
	hipStream_t stream = (hipStream_t)stream_;
	if (rows > VG_NOALLOC_MAX_ROWS)
		return qerr(err, errlen, "%llu rows: the allocation-free path takes at most %u rows", rows,
			VG_NOALLOC_MAX_ROWS);
	if (ncols > VIRG_MAX_COLUMNS || ncols > VG_MAX_COLS - 1)
		return qerr(err, errlen, "%u columns: at most %d", ncols, VIRG_MAX_COLUMNS);
	if (!result)
		return qerr(err, errlen, "no result block");
	int par = -1;
	for (unsigned i = 0; i < num_ops && par < 0; i++)
		if (ops[i].op == OP_Parallel)
			par = (int)i;
	if (par < 0)
		return qerr(err, errlen, "program has no Parallel op");
	virg_tablet_meta tab;
	memset(&tab, 0, sizeof tab);
	tab.fixed_columns = ncols;
	tab.key_type = (virg_t)key_type;
	for (unsigned c = 0; c < ncols; c++) {
		tab.fixed_type[c] = (virg_t)col_types[c];
		tab.fixed_stride[c] = virg_sizeof((virg_t)col_types[c]);
	}
	vg_prog P;
	if (vg_decode(ops, num_ops, (unsigned)par + 1, &tab, nullptr, &P, err, errlen) == VIRG_FAIL)
		return VIRG_FAIL;
	if (vg_kernel_check(&P) != 0)
		return qerr(err, errlen, "op-program needs more register slots than the LDS holds");
	for (unsigned j = 0; j < P.nout; j++)
		if (!out[j])
			return qerr(err, errlen, "result column %u has no output", j);
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) {
		(void)hipGetLastError();
		return qerr(err, errlen, "no current device");
	}
	uint8_t *sp = noalloc_scratch(dev);
	if (!sp)
		return qerr(err, errlen, "scratch of device %d not found", dev);
	vg_tabdesc d;
	memset(&d, 0, sizeof d);
	bool aligned = true;
	for (unsigned c = 0; c < ncols; c++) {
		d.col[c] = (uint64_t)cols[c];
		aligned = aligned && (d.col[c] % 16 == 0);
	}
	d.col[VG_KEY_SRC] = (uint64_t)key;
	aligned = aligned && ((uint64_t)key % 16 == 0);
	d.rows = (uint32_t)rows;
	d.aligned = aligned ? 1u : 0u;
	const uint32_t T = vg_kernel_tile_rows();
	const uint32_t ntiles = (uint32_t)((rows + T - 1) / T);
	vg_ctrl *ctrl = (vg_ctrl *)result;
	if (ntiles == 0) {
		/* nothing to scan: the result block still says 0 rows */
		hipLaunchKernelGGL(vg_noalloc_setup, dim3(1), dim3(256), 0, stream, (vg_tabdesc *)(sp + VG_NA_DESC),
			(uint32_t *)(sp + VG_NA_TILES), 0u, d);
		return herr(hipGetLastError(), "setup") ? VIRG_FAIL : VIRG_SUCCESS;
	}
	hipLaunchKernelGGL(vg_noalloc_setup, dim3((ntiles + 255) / 256), dim3(256), 0, stream,
		(vg_tabdesc *)(sp + VG_NA_DESC), (uint32_t *)(sp + VG_NA_TILES), ntiles, d);
	if (herr(hipGetLastError(), "setup"))
		return VIRG_FAIL;
	vg_launch L;
	memset(&L, 0, sizeof L);
	L.tabs = (const vg_tabdesc *)(sp + VG_NA_DESC);
	L.tile_tab = (const uint32_t *)(sp + VG_NA_TILES);
	L.num_tiles = ntiles;
	L.capacity = capacity;
	L.ctrl = ctrl;
	L.vbits = sp + VG_NA_VBITS;
	L.tcount = (uint32_t *)(sp + VG_NA_TCOUNT);
	L.wgcount = (uint32_t *)(sp + VG_NA_WGCOUNT);
	for (unsigned j = 0; j < P.nout; j++)
		L.out[j] = out[j];
	if (vg_kernel_grid(&P, ntiles) > VG_NOALLOC_MAX_TILES)
		return qerr(err, errlen, "count grid larger than the scratch holds");
	if (vg_kernel_launch(&P, &L, stream, nullptr) != 0)
		return qerr(err, errlen, "kernel launch failed");
	return VIRG_SUCCESS;
}
