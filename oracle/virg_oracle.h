/*
 * ORACLE -- test infrastructure only (see virg_oracle.c).
 */
#ifndef VIRG_ORACLE_H
#define VIRG_ORACLE_H

#include <stdint.h>
#include "virginian.h"

/* Single-threaded reference-order run of the data-parallel section
 * (ops pc .. Converge) over tablets tabs[0..ntab).  Result column j
 * (j < nout) is written to out[j] (up to cap rows; out == NULL counts only).
 * Returns the number of result rows, or -1 with a message in err.
 * *illformed is set when some row read a register it had not written. */
long long oracle_select(const virg_op *stmt, unsigned num_ops, unsigned pc,
	const virg_tablet_meta *const *tabs, unsigned ntab, void **out, unsigned nout,
	unsigned long long cap, int *illformed, char *err, size_t errlen);

/* Multi-threaded run: nthreads threads take whole tablets.  With
 * ordered == 0 rows are reserved per 64-row block under a mutex, as the
 * reference multi-core VM does (row order not deterministic); with
 * ordered == 1 the output equals oracle_select's. */
long long oracle_select_mt(const virg_op *stmt, unsigned num_ops, unsigned pc,
	const virg_tablet_meta *const *tabs, unsigned ntab, void **out, unsigned nout,
	unsigned long long cap, int nthreads, int ordered, int *illformed, char *err, size_t errlen);

#endif
