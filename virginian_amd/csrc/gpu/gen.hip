/*
 * Device table builder: the synthetic tables of host/gen.c, built on the
 * GPU.  The reference's ingest path adds rows one at a time
 * (src/table/insert.c, src/tablet/addrows.c:39-73); a 2e9-row table built
 * that way, or even by 16 host threads, takes about a minute.  Here each
 * tablet's key and six columns are generated by a kernel into device
 * staging and copied into the host tablet (laid out exactly as the host
 * builder lays it out), two tablets in flight so the copy of one overlaps
 * the generation of the next.  The values come from host/gen_math.h, the
 * code the host builder runs, so the tables are identical bit for bit.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../host/internal.h"
#include "../host/gen_math.h"
#include "virg_gpu.h"

/* rows [row0, row0 + n) of the standard columns, one row per thread */
__global__ __launch_bounds__(256) void vg_gen_kernel(uint32_t seed, uint64_t row0, uint32_t n, int64_t *id,
	int32_t *ui, int32_t *ni5, int32_t *ni20, float *uf, float *nf5, float *nf20)
{
	const uint32_t stride = gridDim.x * blockDim.x;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		const uint64_t row = row0 + i;
		id[i] = (int64_t)row;
		ui[i] = vg_gen_uniformi(seed, row);
		ni5[i] = vg_gen_normali(seed, VG_GC_NORMALI5, 5.0, row);
		ni20[i] = vg_gen_normali(seed, VG_GC_NORMALI20, 20.0, row);
		uf[i] = vg_gen_uniformf(seed, row);
		nf5[i] = vg_gen_normalf(seed, VG_GC_NORMALF5, 5.0, row);
		nf20[i] = vg_gen_normalf(seed, VG_GC_NORMALF20, 20.0, row);
	}
}

extern "C" int vg_gen_table(virginian *v, const char *name, unsigned long long rows, unsigned tablet_rows,
	unsigned seed)
{
	virg_state *s = virg_state_of(v);
	VIRG_CHECK(s == NULL || s->dev == NULL, "vg_gen_table needs a HIP device");
	VIRG_CHECK(tablet_rows == 0, "vg_gen_table builds whole tablets: tablet_rows must be > 0");
	unsigned tid;
	VIRG_CHECK(virg_gen_schema(v, name, &tid) == VIRG_FAIL, "table create failed");
	const unsigned cap = (tablet_rows + 15u) & ~15u;
	const size_t tr = tablet_rows;
	char *buf[2] = { nullptr, nullptr };
	hipStream_t st[2] = { nullptr, nullptr };
	hipEvent_t done[2] = { nullptr, nullptr };
	int ok = 1;
	for (int b = 0; b < 2 && ok; b++)
		ok = hipMalloc((void **)&buf[b], tr * 32) == hipSuccess &&
			hipStreamCreateWithFlags(&st[b], hipStreamNonBlocking) == hipSuccess &&
			hipEventCreateWithFlags(&done[b], hipEventDisableTiming) == hipSuccess;
	/* VIRG_GEN_PROGRESS=n: a line on stderr every n rows (long tables) */
	const char *pe = getenv("VIRG_GEN_PROGRESS");
	const uint64_t every = pe ? strtoull(pe, NULL, 10) : 0;
	uint64_t k = 0;
	for (uint64_t r0 = 0; ok && r0 < rows; r0 += tr, k++) {
		const int b = (int)(k & 1);
		const unsigned n = (unsigned)(rows - r0 < tr ? rows - r0 : tr);
		/* this staging buffer's copies of two tablets ago have landed */
		if (k >= 2 && hipEventSynchronize(done[b]) != hipSuccess) {
			ok = 0;
			break;
		}
		if (virg_table_appendtablet(v, tid, cap, n, NULL, NULL) == VIRG_FAIL) {
			ok = 0;
			break;
		}
		const unsigned tabid = v->db.last_tablet[tid];
		virg_tablet_meta *m = NULL;
		if (virg_db_load(v, tabid, &m) == VIRG_FAIL) {
			ok = 0;
			break;
		}
		virg_tablet_unlock(v, tabid);
		if (m->key_stride != 8 || m->fixed_columns != 6) {
			fprintf(stderr, "vg_gen_table: unexpected tablet layout\n");
			ok = 0;
			break;
		}
		int64_t *d_id = (int64_t *)buf[b];
		char *d_col[6];
		for (int j = 0; j < 6; j++)
			d_col[j] = buf[b] + 8 * tr + 4 * tr * (size_t)j;
		const unsigned grid = (n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192;
		hipLaunchKernelGGL(vg_gen_kernel, dim3(grid), dim3(256), 0, st[b], (uint32_t)seed, (uint64_t)r0, n, d_id,
			(int32_t *)d_col[0], (int32_t *)d_col[1], (int32_t *)d_col[2], (float *)d_col[3],
			(float *)d_col[4], (float *)d_col[5]);
		ok = hipGetLastError() == hipSuccess &&
			hipMemcpyAsync((char *)m + m->key_block, d_id, 8 * (size_t)n, hipMemcpyDeviceToHost, st[b]) ==
				hipSuccess;
		for (int j = 0; ok && j < 6; j++)
			ok = m->fixed_stride[j] == 4 &&
				hipMemcpyAsync((char *)m + m->fixed_block + m->fixed_offset[j], d_col[j], 4 * (size_t)n,
					hipMemcpyDeviceToHost, st[b]) == hipSuccess;
		ok = ok && hipEventRecord(done[b], st[b]) == hipSuccess;
		if (ok && every && (r0 + n) / every != r0 / every) {
			fprintf(stderr, "# generated %llu of %llu rows on the device\n", (unsigned long long)(r0 + n),
				(unsigned long long)rows);
			fflush(stderr);
		}
	}
	for (int b = 0; b < 2; b++) {
		if (st[b] && hipStreamSynchronize(st[b]) != hipSuccess)
			ok = 0;
		if (done[b])
			(void)hipEventDestroy(done[b]);
		if (st[b])
			(void)hipStreamDestroy(st[b]);
		if (buf[b])
			(void)hipFree(buf[b]);
	}
	if (!ok)
		(void)hipGetLastError();
	VIRG_CHECK(!ok, "device table generation failed");
	return VIRG_SUCCESS;
}
