/*
 * Serving loop in the host library: requests (op-programs) arrive over
 * time and are run against resident tables with continuous batching.
 * No counterpart in the reference, whose queries run one at a time
 * (src/vm/execute.c:28-102).
 *
 * One host thread drives the loop (vg_server_run, or vg_server_step from
 * the caller's own loop).  Each pass:
 *   1. requests whose arrival time has come join the waiting line;
 *   2. finished work is collected: a lone run when its write pass has
 *      stored the outcome into pinned host memory (polled, no runtime
 *      call), a batch when its stream is idle (the kernels stored every
 *      query's outcome into pinned memory too); a query whose result did
 *      not fit its buffers goes back to the front of the line with its
 *      exact size, the others are complete (a kept result is copied to
 *      host memory first) and their buffers go back to the arena;
 *   3. while a runner (a stream) is free, waiting queries are admitted in
 *      arrival order (reruns first) as long as their result buffers fit in
 *      the arena: admission by HBM footprint.  While the line is short a
 *      run takes one query, alone as its own count and write launches, and
 *      at most one costly query (rows x ops) runs at a time -- the
 *      interpreter is bound by its scalar dispatch, so two costly runs side
 *      by side take as long as one after the other while each is slowed,
 *      whereas cheap queries beside a costly one pass it without slowing it
 *      much; once the line is long, runs take many queries as one
 *      persistent batch (vg_pbatch).  Runs on different runners' streams
 *      overlap on the device.
 *
 * Buffers are sized for every row of the query's table, by the
 * independence estimate of the filter over the table's histograms, or by
 * the result-size network (per tablet, plus a margin); the exact count of
 * a run that overflowed sizes its rerun, so results never depend on the
 * estimate.  With one runner and one query per batch the loop is the plain
 * way: each request alone, in arrival order, waited for, then the next.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <deque>
#include <map>
#include <unordered_map>
#include <vector>

#include "virg_gpu.h"
#include "vg_prog.h"

extern "C" int vgd_query_launch_lone(vg_query *q, hipStream_t stream, vg_bresult *res);
extern "C" hipStream_t vgd_pbatch_stream(vg_pbatch *b);
extern "C" const vg_bresult *vgd_pbatch_outcomes(vg_pbatch *b);

namespace {

double mono_now()
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

size_t round256(size_t x)
{
	return (x + 255) & ~(size_t)255;
}

/* First fit over one block of device memory; free blocks by offset,
 * neighbours merged on release. */
struct arena {
	char *base = nullptr;
	size_t size = 0, used = 0, peak = 0;
	std::map<size_t, size_t> free_;      /* offset -> bytes */

	bool alloc(size_t n, size_t *off)
	{
		n = round256(n);
		for (auto it = free_.begin(); it != free_.end(); ++it) {
			if (it->second < n)
				continue;
			*off = it->first;
			const size_t rest = it->second - n;
			const size_t at = it->first + n;
			free_.erase(it);
			if (rest)
				free_[at] = rest;
			used += n;
			peak = used > peak ? used : peak;
			return true;
		}
		return false;
	}
	void release(size_t off, size_t n)
	{
		n = round256(n);
		used -= n;
		auto it = free_.emplace(off, n).first;
		auto nx = std::next(it);
		if (nx != free_.end() && it->first + it->second == nx->first) {
			it->second += nx->second;
			free_.erase(nx);
		}
		if (it != free_.begin()) {
			auto pv = std::prev(it);
			if (pv->first + pv->second == it->first) {
				pv->second += it->second;
				free_.erase(it);
			}
		}
	}
};

/* a registered table's statistics (owned copies) */
struct table_info {
	unsigned ntab = 0;
	std::vector<double> rows, edges, hist, tabfeat;
	std::vector<int> is_int;
	vg_table_stats view() const
	{
		vg_table_stats st;
		st.ntab = ntab;
		st.rows = rows.data();
		st.edges = edges.data();
		st.hist = hist.data();
		st.is_int = is_int.data();
		st.tabfeat = tabfeat.data();
		return st;
	}
};

enum { R_PENDING = 0, R_WAITING, R_RUNNING, R_DONE, R_FAILED };

struct request {
	std::vector<virg_op> ops;
	unsigned table = 0;
	double arrival = 0, ready = -1;
	int state = R_PENDING;
	unsigned reruns = 0;
	uint64_t cap = 0;           /* result rows reserved (0: not sized yet) */
	uint64_t exact = 0;         /* rerun size (0: none) */
	uint64_t rows = 0;
	size_t need = 0, off = 0;
	bool holds = false;         /* owns [off, off + need) of the arena */
	vg_query *q = nullptr;
	std::vector<unsigned> widths;
	uint64_t table_rows = 0;
	uint64_t cost = 0;          /* rows x ops: how long the interpreter runs it */
	double seen = -1;           /* when its outcome was seen complete in a running batch */
	std::vector<std::vector<char>> host;   /* kept results, one array per column */
};

struct runner {
	vg_pbatch *pb = nullptr;
	hipStream_t stream = nullptr;
	vg_bresult *res = nullptr;  /* lone run's outcome (pinned, coherent) */
	int busy = 0;               /* 0 free, 1 lone run, 2 batch */
	bool costly = false;        /* a lone run of a costly request */
	std::vector<size_t> reqs;
};

} // namespace

struct vg_server {
	virginian *v = nullptr;
	vg_server_config cfg;
	arena ar;
	std::unordered_map<unsigned, table_info> tables;
	std::vector<request> reqs;
	std::deque<size_t> waiting;
	std::deque<size_t> reruns;
	size_t next_arrival = 0;    /* requests before it have arrived (arrival order) */
	std::vector<size_t> by_arrival;
	bool order_dirty = false;
	std::vector<runner> runners;
	hipStream_t copy = nullptr;  /* kept results come back on it */
	char *hbase = nullptr;       /* host_arena: the host's view of the arena (ar.base is the device's) */
	double t0 = 0;
	size_t done = 0, failed = 0;
	vg_server_stats st;
	std::vector<float> feat, pred;
	char msg[256] = {0};
};

static int serr(vg_server *s, char *err, size_t n, const char *what)
{
	if (err && n)
		snprintf(err, n, "%s", what);
	if (s)
		snprintf(s->msg, sizeof s->msg, "%s", what);
	return VIRG_FAIL;
}

extern "C" void vg_server_free(vg_server *s)
{
	if (!s)
		return;
	for (runner &r : s->runners) {
		if (r.stream)
			(void)hipStreamSynchronize(r.stream);
		if (r.pb) {
			char e[64];
			(void)vg_pbatch_reload(r.pb, nullptr, 0, e, sizeof e);
			vg_pbatch_free(r.pb);
		}
		if (r.res)
			(void)hipHostFree(r.res);
	}
	for (request &q : s->reqs)
		if (q.q)
			vg_query_free(q.q);
	if (s->copy)
		(void)hipStreamDestroy(s->copy);
	if (s->hbase)
		(void)hipHostFree(s->hbase);
	else if (s->ar.base)
		vg_device_free(s->v, s->ar.base);
	delete s;
}

extern "C" vg_server *vg_server_create(virginian *v, const vg_server_config *cfg, char *err, size_t errlen)
{
	vg_server *s = new vg_server();
	s->v = v;
	memset(&s->cfg, 0, sizeof s->cfg);
	if (cfg)
		s->cfg = *cfg;
	vg_server_config &c = s->cfg;
	if (!c.arena_bytes)
		c.arena_bytes = 8ull << 30;
	if (!c.runners)
		c.runners = 3;
	if (!c.max_batch)
		c.max_batch = 64;
	if (!c.tiles_per_item)
		c.tiles_per_item = 8;
	if (!c.batch_line)
		c.batch_line = 16;
	if (!c.costly)
		c.costly = 100000000ull;
	memset(&s->st, 0, sizeof s->st);
	if (c.sizing > VG_SIZE_NET || (c.sizing == VG_SIZE_NET && !c.net)) {
		serr(s, err, errlen, "vg_server_create: sizing by network needs the network");
		delete s;
		return nullptr;
	}
	if (c.host_arena) {
		/* results written by the kernels straight into pinned host memory */
		if (hipHostMalloc((void **)&s->hbase, c.arena_bytes, hipHostMallocMapped | hipHostMallocCoherent) !=
				hipSuccess ||
			hipHostGetDevicePointer((void **)&s->ar.base, s->hbase, 0) != hipSuccess) {
			(void)hipGetLastError();
			if (s->hbase)
				(void)hipHostFree(s->hbase);
			s->hbase = nullptr;
			s->ar.base = nullptr;
		}
	} else {
		s->ar.base = (char *)vg_device_alloc(v, c.arena_bytes);
	}
	if (!s->ar.base) {
		serr(s, err, errlen, "vg_server_create: cannot allocate the result arena");
		delete s;
		return nullptr;
	}
	s->ar.size = c.arena_bytes;
	s->ar.free_[0] = c.arena_bytes;
	if (c.keep && !c.host_arena && hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess) {
		(void)hipGetLastError();
		vg_server_free(s);
		serr(nullptr, err, errlen, "vg_server_create: cannot create a stream");
		return nullptr;
	}
	for (unsigned i = 0; i < c.runners; i++) {
		runner r;
		vg_pbatch_config pc;
		memset(&pc, 0, sizeof pc);
		pc.tiles_per_item = c.tiles_per_item;
		pc.grid_div = c.grid_div;
		pc.split = c.split;
		pc.order = 1;
		pc.direct = 1;
		pc.db = v;
		pc.single = c.single;
		char e[256];
		r.pb = vg_pbatch_create(nullptr, 0, &pc, e, sizeof e);
		if (!r.pb || hipHostMalloc((void **)&r.res, sizeof(vg_bresult), hipHostMallocCoherent) != hipSuccess) {
			if (r.pb)
				vg_pbatch_free(r.pb);
			(void)hipGetLastError();
			vg_server_free(s);
			serr(nullptr, err, errlen, "vg_server_create: cannot set up a runner");
			return nullptr;
		}
		r.stream = vgd_pbatch_stream(r.pb);
		s->runners.push_back(r);
	}
	s->t0 = mono_now();
	return s;
}

extern "C" int vg_server_table(vg_server *s, unsigned table_id, const vg_table_stats *st)
{
	if (!st || st->ntab == 0)
		return serr(s, nullptr, 0, "vg_server_table: no statistics");
	table_info t;
	const size_t nt = st->ntab, nc = VG_FEAT_COLS, nb = VG_FEAT_BINS;
	t.ntab = st->ntab;
	t.rows.assign(st->rows, st->rows + nt);
	t.edges.assign(st->edges, st->edges + nc * (nb + 1));
	t.hist.assign(st->hist, st->hist + nt * nc * nb);
	t.is_int.assign(st->is_int, st->is_int + nc);
	t.tabfeat.assign(st->tabfeat, st->tabfeat + nt * nc * (nb + 2));
	s->tables[table_id] = std::move(t);
	return VIRG_SUCCESS;
}

extern "C" double vg_server_now(vg_server *s)
{
	return mono_now() - s->t0;
}

extern "C" int vg_server_submit(vg_server *s, const virg_op *ops, unsigned num_ops, double arrival,
	unsigned long long *id)
{
	if (!ops || num_ops == 0 || ops[0].op != OP_Table)
		return serr(s, nullptr, 0, "vg_server_submit: the program must start with its Table op");
	request r;
	r.ops.assign(ops, ops + num_ops);
	r.table = (unsigned)ops[0].p1;
	r.arrival = arrival < 0 ? vg_server_now(s) : arrival;
	for (unsigned i = 0; i < num_ops; i++)
		if (ops[i].op == OP_ResultColumn) {
			r.widths.push_back(virg_sizes[ops[i].p1]);   /* p1: the column type */
			r.ops[i].p4.li = 0;     /* the column's name: the caller's memory, not kept */
		}
	if (!s->by_arrival.empty() && r.arrival < s->reqs[s->by_arrival.back()].arrival)
		s->order_dirty = true;
	s->by_arrival.push_back(s->reqs.size());
	if (id)
		*id = s->reqs.size();
	s->reqs.push_back(std::move(r));
	return VIRG_SUCCESS;
}

/* Result rows to reserve for r: every row, or per tablet
 * rows * min(1, max(p, 0) + margin) (f32, as the sizing was calibrated)
 * summed, plus 64. */
static int size_request(vg_server *s, request &r)
{
	if (r.exact) {
		r.cap = r.exact;
		return VIRG_SUCCESS;
	}
	if (s->cfg.sizing == VG_SIZE_ALL) {
		r.cap = r.table_rows;
		return VIRG_SUCCESS;
	}
	auto it = s->tables.find(r.table);
	if (it == s->tables.end())
		return serr(s, nullptr, 0, "vg_server: sizing needs the statistics of the query's table");
	const vg_table_stats st = it->second.view();
	s->feat.resize((size_t)st.ntab * VG_FEATURES);
	s->pred.resize(st.ntab);
	if (vg_program_features(r.ops.data(), (unsigned)r.ops.size(), &st, s->feat.data()) != VIRG_SUCCESS)
		return serr(s, nullptr, 0, "vg_server: features failed");
	if (s->cfg.sizing == VG_SIZE_NET) {
		if (vg_mlp_forward_host(s->cfg.net, s->feat.data(), st.ntab, s->pred.data()) != VIRG_SUCCESS)
			return serr(s, nullptr, 0, "vg_server: predictor failed");
	} else {
		for (unsigned t = 0; t < st.ntab; t++)
			s->pred[t] = s->feat[(size_t)t * VG_FEATURES + 3];
	}
	const float m = s->cfg.margin;
	double sum = 0;
	for (unsigned t = 0; t < st.ntab; t++) {
		float p = s->pred[t] > 0.f ? s->pred[t] : 0.f;
		p = p + m;
		p = p < 1.f ? p : 1.f;
		sum += st.rows[t] * (double)p;
	}
	r.cap = (uint64_t)ceil(sum) + 64;
	return VIRG_SUCCESS;
}

/* Prepare r's query (once) and its size; false when it cannot run. */
static bool prepare(vg_server *s, request &r)
{
	if (!r.q) {
		char e[256];
		r.q = vg_query_create(s->v, r.ops.data(), (unsigned)r.ops.size(), 1,
			VG_QUERY_LIGHT | VG_QUERY_NO_OUTPUT | VG_QUERY_INTERP, e, sizeof e);
		if (!r.q) {
			serr(s, nullptr, 0, e);
			return false;
		}
		unsigned long long rows = 0;
		unsigned tiles, rb, wb, nout;
		vg_query_info(r.q, &rows, &tiles, &rb, &wb, &nout);
		r.table_rows = rows;
		r.cost = rows * (uint64_t)r.ops.size();
	}
	if (r.cap == 0) {
		if (size_request(s, r) != VIRG_SUCCESS)
			return false;
		uint64_t c = r.cap ? r.cap : 1;
		r.need = 0;
		for (unsigned w : r.widths)
			r.need += round256(c * w);
		s->st.reserved_bytes += r.need;
	}
	return true;
}

static void fail_request(vg_server *s, request &r)
{
	r.state = R_FAILED;
	r.ready = vg_server_now(s);
	s->failed++;
	if (r.q) {
		vg_query_free(r.q);
		r.q = nullptr;
	}
}

/* Give r its buffers (false: the arena has no room now). */
static bool place(vg_server *s, request &r)
{
	size_t off;
	if (!s->ar.alloc(r.need, &off))
		return false;
	r.off = off;
	r.holds = true;
	void *outs[VG_MAX_OUT];
	size_t o = off;
	const uint64_t c = r.cap ? r.cap : 1;
	for (size_t j = 0; j < r.widths.size(); j++) {
		outs[j] = s->ar.base + o;
		o += round256(c * r.widths[j]);
	}
	if (vg_query_set_outputs(r.q, outs, c) != VIRG_SUCCESS) {
		s->ar.release(r.off, r.need);
		r.holds = false;
		return false;
	}
	return true;
}

static void release_buffers(vg_server *s, request &r)
{
	if (r.holds) {
		s->ar.release(r.off, r.need);
		r.holds = false;
	}
}

/* A finished run of r: complete, or back to the line at its exact size. */
static void finish(vg_server *s, request &r, uint64_t total, unsigned flags)
{
	if (flags) {
		/* a single-pass wait that gave up leaves no usable count */
		r.exact = (flags & VG_OVF_WAIT) ? r.table_rows : total;
		if (r.exact == 0)
			r.exact = 1;
		r.cap = 0;
		r.seen = -1;
		r.reruns++;
		s->st.reruns++;
		release_buffers(s, r);
		r.state = R_WAITING;
		s->reruns.push_back((size_t)(&r - s->reqs.data()));
		return;
	}
	r.rows = total;
	if (s->cfg.keep) {
		/* the result to host memory; the buffers go back at once */
		r.host.assign(r.widths.size(), std::vector<char>());
		size_t o = r.off;
		const uint64_t c = r.cap ? r.cap : 1;
		bool copied = true;
		for (size_t j = 0; j < r.widths.size(); j++) {
			r.host[j].resize(total * r.widths[j]);
			if (s->hbase)               /* already in host memory: no copy engine */
				memcpy(r.host[j].data(), s->hbase + o, total * r.widths[j]);
			else if (total && hipMemcpyAsync(r.host[j].data(), s->ar.base + o, total * r.widths[j],
					hipMemcpyDeviceToHost, s->copy) != hipSuccess)
				copied = false;
			o += round256(c * r.widths[j]);
		}
		if (!s->hbase && hipStreamSynchronize(s->copy) != hipSuccess)
			copied = false;
		if (!copied) {
			/* no result to hand over: the request fails, it is not done */
			(void)hipGetLastError();
			std::vector<std::vector<char>>().swap(r.host);
			snprintf(s->msg, sizeof s->msg, "a kept result could not be copied to host memory");
			release_buffers(s, r);
			fail_request(s, r);
			return;
		}
	}
	r.state = R_DONE;
	/* complete in HBM when the batch stored its outcome (kept results:
	 * once copied out) */
	r.ready = !s->cfg.keep && r.seen >= 0 ? r.seen : vg_server_now(s);
	s->done++;
	vg_query_free(r.q);
	r.q = nullptr;
	release_buffers(s, r);
}

static int collect(vg_server *s)
{
	int completed = 0;
	for (runner &rn : s->runners) {
		if (rn.busy == 1) {
			if (__atomic_load_n(&rn.res->done, __ATOMIC_ACQUIRE) == 0)
				continue;
			const double ta = mono_now();
			request &r = s->reqs[rn.reqs[0]];
			const size_t before = s->done;
			finish(s, r, rn.res->total, rn.res->flags);
			completed += (int)(s->done - before);
			rn.busy = 0;
			rn.reqs.clear();
			s->st.collect_s += mono_now() - ta;
		} else if (rn.busy == 2) {
			if (!vg_pbatch_done(rn.pb)) {
				/* the single pass stores each query's outcome when its
				 * last chunk is written: note complete results early */
				const vg_bresult *o = vgd_pbatch_outcomes(rn.pb);
				double now = -1;
				for (size_t i = 0; o && i < rn.reqs.size(); i++) {
					request &r = s->reqs[rn.reqs[i]];
					if (r.seen >= 0 || __atomic_load_n(&o[i].done, __ATOMIC_ACQUIRE) == 0 || o[i].flags)
						continue;
					if (now < 0)
						now = vg_server_now(s);
					r.seen = now;
				}
				continue;
			}
			const double ta = mono_now();
			const size_t n = rn.reqs.size();
			std::vector<unsigned long long> rows(n);
			std::vector<unsigned> flags(n);
			if (vg_pbatch_results(rn.pb, rows.data(), nullptr, flags.data()) != VIRG_SUCCESS)
				return -1;
			char e[64];
			/* the runner lets go of its queries before any is freed */
			if (vg_pbatch_reload(rn.pb, nullptr, 0, e, sizeof e) != VIRG_SUCCESS)
				return -1;
			const size_t before = s->done;
			for (size_t i = 0; i < n; i++)
				finish(s, s->reqs[rn.reqs[i]], rows[i], flags[i]);
			completed += (int)(s->done - before);
			rn.busy = 0;
			rn.reqs.clear();
			s->st.collect_s += mono_now() - ta;
		}
	}
	return completed;
}

/* Admit from `line` onto runner rn, in arrival order, at most max_take
 * requests; a costly request only when allow_costly (those left wait in
 * place, cheaper ones behind them may go first).  *blocked = the arena had
 * no room for the next one.  False when nothing was admitted. */
static bool admit(vg_server *s, std::deque<size_t> &line, runner &rn, size_t max_take, bool allow_costly,
	bool *blocked)
{
	const double ta = mono_now();
	std::vector<size_t> take;
	*blocked = false;
	for (size_t i = 0; i < line.size() && take.size() < max_take;) {
		request &r = s->reqs[line[i]];
		if (!prepare(s, r)) {
			fail_request(s, r);
			line.erase(line.begin() + (long)i);
			continue;
		}
		if (r.need > s->ar.size) {
			snprintf(s->msg, sizeof s->msg, "a request's result buffers (%zu bytes) exceed the arena", r.need);
			fail_request(s, r);
			line.erase(line.begin() + (long)i);
			continue;
		}
		if (!allow_costly && r.cost >= s->cfg.costly) {
			++i;
			continue;
		}
		if (!place(s, r)) {
			*blocked = true;
			break;                  /* admission is in arrival order */
		}
		take.push_back(line[i]);
		line.erase(line.begin() + (long)i);
	}
	const double tb = mono_now();
	s->st.admit_s += tb - ta;
	if (take.empty())
		return false;
	bool ok;
	if (take.size() == 1 && !s->cfg.single) {
		request &r = s->reqs[take[0]];
		ok = vgd_query_launch_lone(r.q, rn.stream, rn.res) == VIRG_SUCCESS;
		rn.busy = 1;
		rn.costly = r.cost >= s->cfg.costly;
		s->st.lone++;
	} else {
		std::vector<vg_query *> qs;
		for (size_t i : take)
			qs.push_back(s->reqs[i].q);
		char e[256];
		ok = vg_pbatch_reload(rn.pb, qs.data(), (unsigned)qs.size(), e, sizeof e) == VIRG_SUCCESS;
		/* the last run's outcomes stay until this run's reset clears them:
		 * clear them first, the runner being idle */
		vg_bresult *o = (vg_bresult *)vgd_pbatch_outcomes(rn.pb);
		if (ok && o)
			memset(o, 0, qs.size() * sizeof(vg_bresult));
		ok = ok && vg_pbatch_launch(rn.pb) == VIRG_SUCCESS;
		if (!ok)
			snprintf(s->msg, sizeof s->msg, "%s", e);
		rn.busy = 2;
		s->st.batches++;
	}
	for (size_t i : take)
		s->reqs[i].state = R_RUNNING;
	rn.reqs = take;
	s->st.launch_s += mono_now() - tb;
	s->st.in_flight_peak = s->ar.used > s->st.in_flight_peak ? s->ar.used : s->st.in_flight_peak;
	return ok;
}

static void sort_arrivals(vg_server *s)
{
	if (!s->order_dirty)
		return;
	std::stable_sort(s->by_arrival.begin() + (long)s->next_arrival, s->by_arrival.end(),
		[s](size_t a, size_t b) { return s->reqs[a].arrival < s->reqs[b].arrival; });
	s->order_dirty = false;
}

extern "C" int vg_server_step(vg_server *s)
{
	sort_arrivals(s);
	const double now = vg_server_now(s);
	while (s->next_arrival < s->by_arrival.size() && s->reqs[s->by_arrival[s->next_arrival]].arrival <= now) {
		request &r = s->reqs[s->by_arrival[s->next_arrival++]];
		if (r.state == R_PENDING) {
			r.state = R_WAITING;
			s->waiting.push_back((size_t)(&r - s->reqs.data()));
		}
	}
	const int completed = collect(s);
	if (completed < 0)
		return -1;
	for (runner &rn : s->runners) {
		if (rn.busy)
			continue;
		/* reruns first (they arrived earlier), then new requests */
		std::deque<size_t> &line = !s->reruns.empty() ? s->reruns : s->waiting;
		if (line.empty())
			break;
		bool busy_any = false, costly_running = false;
		for (runner &o : s->runners) {
			busy_any = busy_any || o.busy;
			costly_running = costly_running || (o.busy == 1 && o.costly);
		}
		/* a short line: runs of one request each, at most one costly run
		 * at a time (costly runs side by side only slow each other); a
		 * line of batch_line or more: batches */
		const bool light = line.size() < s->cfg.batch_line;
		bool blocked = false;
		if (!admit(s, line, rn, light ? 1 : s->cfg.max_batch, !light || !costly_running, &blocked)) {
			if (rn.busy)
				return -1;      /* a launch failed */
			if (!blocked)
				break;          /* only costly requests wait, behind a costly run */
			if (!busy_any && !line.empty() && s->reqs[line.front()].state == R_WAITING) {
				/* nothing runs and the head does not fit: it never will */
				request &r = s->reqs[line.front()];
				release_buffers(s, r);
				snprintf(s->msg, sizeof s->msg, "a request's result buffers do not fit in the arena");
				fail_request(s, r);
				line.pop_front();
			}
			break;
		}
	}
	return completed;
}

extern "C" int vg_server_run(vg_server *s, double timeout_s)
{
	s->t0 = mono_now();
	const size_t n = s->reqs.size();
	while (s->done + s->failed < n) {
		if (vg_server_step(s) < 0)
			return VIRG_FAIL;
		if (timeout_s > 0 && vg_server_now(s) > timeout_s)
			return serr(s, nullptr, 0, "vg_server_run: timed out");
	}
	for (runner &rn : s->runners)
		if (rn.stream && hipStreamSynchronize(rn.stream) != hipSuccess)
			return serr(s, nullptr, 0, "vg_server_run: a run failed");
	return VIRG_SUCCESS;
}

extern "C" int vg_server_request(vg_server *s, unsigned long long id, vg_server_req *out)
{
	if (id >= s->reqs.size())
		return VIRG_FAIL;
	const request &r = s->reqs[id];
	out->state = (unsigned)r.state;
	out->reruns = r.reruns;
	out->rows = r.rows;
	out->capacity = r.cap;
	out->arrival = r.arrival;
	out->ready = r.ready;
	return VIRG_SUCCESS;
}

extern "C" int vg_server_fetch(vg_server *s, unsigned long long id, void **cols, unsigned long long max_rows)
{
	if (id >= s->reqs.size() || s->reqs[id].state != R_DONE || s->reqs[id].host.size() != s->reqs[id].widths.size())
		return serr(s, nullptr, 0, "vg_server_fetch: no kept result for this request");
	const request &r = s->reqs[id];
	const uint64_t n = r.rows < max_rows ? r.rows : max_rows;
	for (size_t j = 0; j < r.widths.size(); j++)
		if (cols[j] && n)
			memcpy(cols[j], r.host[j].data(), n * r.widths[j]);
	return VIRG_SUCCESS;
}

extern "C" int vg_server_release(vg_server *s, unsigned long long id)
{
	if (id >= s->reqs.size() || s->reqs[id].state != R_DONE)
		return VIRG_FAIL;
	std::vector<std::vector<char>>().swap(s->reqs[id].host);
	return VIRG_SUCCESS;
}

extern "C" int vg_server_get_stats(vg_server *s, vg_server_stats *out)
{
	*out = s->st;
	out->done = s->done;
	out->failed = s->failed;
	out->arena_peak = s->ar.peak;
	return VIRG_SUCCESS;
}

extern "C" const char *vg_server_error(vg_server *s)
{
	return s->msg;
}
