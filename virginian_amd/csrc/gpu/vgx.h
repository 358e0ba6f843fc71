/*
 * Prepare-time specialised kernels (jit.hip): one single-pass kernel per
 * decoded program structure, compiled with hiprtc and cached.
 */
#ifndef VGX_H
#define VGX_H

#include <stdint.h>
#include <hip/hip_runtime.h>
#include "vg_prog.h"
#include "vgx_abi.h"

/* geometry of a specialised kernel */
typedef struct {
	int R;                     /* rows per lane per tile (4, 8 or 16) */
	uint32_t S;                /* tiles per chunk (one grab, one look-back) */
	uint32_t TR;               /* rows per tile = 256 * R */
	uint32_t fbytes_row;       /* bytes per row of the filter columns */
	uint32_t ftile_bytes;      /* LDS bytes of one tile's staged filter columns */
	uint32_t stage_off;        /* LDS offset of the output staging area */
	uint32_t stage_wave_bytes; /* output staging bytes per wave */
	uint32_t lds_bytes;        /* dynamic LDS per workgroup */
	int aligned;               /* every tablet's column bases are 16-byte aligned:
	                            * result columns load 16 bytes per lane */
	int staged;                /* 1: a chunk's filter columns are staged in LDS in
	                            * the filter phase and read again from there;
	                            * 0: loaded into registers tile by tile, and those
	                            * the result phase needs loaded again like the
	                            * result columns (for the rows selected) */
	int pipe;                  /* 1: the kernel counts its next chunk before it looks
	                            * back for the current one (two staging buffers) */
	int held;                  /* 1: the filter's one column is loaded into registers
	                            * once per chunk and held there for both phases (no
	                            * LDS staging, no second read) */
	int waves;                 /* minimum waves per SIMD the kernel is compiled for
	                            * (its launch bound), 0 = the compiler's choice */
} vgx_shape;

typedef struct vgx_kernel vgx_kernel;

/* shape for a program (R_override, S_override: 0 = chosen from the program;
 * pipe 1: for the one-query kernel that overlaps a chunk's look-back with
 * the next chunk's filter phase; 2: a batch's chunk function with the
 * pipelined kernel's smaller chunks, not pipelined) */
int vgx_shape_of(const vg_prog *P, int aligned, int R_override, int S_override, int pipe, vgx_shape *sh,
	char *err, size_t errlen);
/* the `pipe` argument of a batch's shapes */
int vgxb_shape_mode(void);
/* the immediates of P as its kernel reads them (a->imm, a->oimm) */
void vgx_immediates(const vg_prog *P, vgx_launch_args *a);
/* kernel source (malloc'd) */
int vgx_source(const vg_prog *P, const vgx_shape *sh, char **src);
/* compiled kernel for P (from the cache, or compiled now); NULL + message */
vgx_kernel *vgx_get(const vg_prog *P, int aligned, char *err, size_t errlen);
const vgx_shape *vgx_kernel_shape(const vgx_kernel *k);
double vgx_kernel_compile_ms(const vgx_kernel *k);
/* workgroups for a launch of nchunks chunks on `cus` compute units */
unsigned vgx_kernel_grid(const vgx_kernel *k, unsigned nchunks, unsigned cus);

/* a batch module for programs P[0..n) (aligned[i]: query i's tablets are
 * 16-byte aligned); fn[i] = query i's program index in it */
vgx_kernel *vgxb_get(const vg_prog *const *P, const int *aligned, unsigned n, uint32_t *fn, char *err,
	size_t errlen);
unsigned vgxb_grid(const vgx_kernel *k, unsigned nitems, unsigned cus);

int vgx_launch(const vgx_kernel *k, const vgx_launch_args *a, unsigned grid, hipStream_t stream);
int vgxb_launch(const vgx_kernel *k, const vgxb_args *a, unsigned grid, hipStream_t stream);

#endif
