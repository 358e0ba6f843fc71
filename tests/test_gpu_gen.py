"""The device table builder (vg_gen_table) against the host generator,
which the known-answer and distribution tests pin: for several sizes,
tablet sizes (ragged last tablets included) and seeds, every tablet's
rows, key and six columns are identical bit for bit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import virginian_amd as va

pytestmark = pytest.mark.gpu

CASES = [(1, 16, 3), (1000, 1000, 1), (250_001, 65_536, 9), (3_000_017, 1 << 20, 2)]


def _columns(db, tid) -> list:
    """Per tablet: (rows, key, six columns) over the rows it holds."""
    out = []
    for t in db.tablets(tid):
        m = t.contents
        base = C.addressof(m)
        n = m.rows
        key = np.ctypeslib.as_array((C.c_int64 * n).from_address(base + m.key_block)).copy()
        cols = [np.ctypeslib.as_array((C.c_uint32 * n).from_address(base + m.fixed_block + m.fixed_offset[j])).copy()
                for j in range(m.fixed_columns)]
        out.append((n, key, cols))
    return out


@pytest.mark.parametrize("rows,tablet_rows,seed", CASES)
def test_device_builder_matches_host(gpu_db, rows, tablet_rows, seed):
    print(f"gen rows={rows} tablet_rows={tablet_rows} seed={seed}", flush=True)
    host = gpu_db.gen_table("h", rows, tablet_rows, seed=seed)
    dev = gpu_db.gen_table("d", rows, tablet_rows, seed=seed, device=True)
    a, b = _columns(gpu_db, host), _columns(gpu_db, dev)
    assert len(a) == len(b) == -(-rows // tablet_rows)
    assert sum(n for n, _, _ in b) == rows
    for k, ((na, ka, ca), (nb, kb, cb)) in enumerate(zip(a, b)):
        assert na == nb, k
        assert np.array_equal(ka, kb), k
        for j in range(6):
            assert np.array_equal(ca[j], cb[j]), (k, j)
