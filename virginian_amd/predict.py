"""Learned result sizing for single-pass execution.

A query run in one pass (vg_pbatch single mode) writes its rows straight
into output buffers that must be sized before it runs.  The predictor
estimates, per tablet, the fraction of rows a program selects; the buffer
is the sum over tablets of rows * min(1, prediction + margin).  A run whose
result does not fit reports it and is run again by the exact two-pass
executor, so results never depend on the prediction: a better predictor
only means smaller buffers (more queries admitted into HBM) and fewer
reruns.

Features of a (program, tablet) pair:
  - the filter evaluated under independence: each `column op constant`
    compare passing at its histogram estimate on the tablet, and any other
    compare at 1/2 (the estimate the network corrects) -- and in six more
    variants with those compares at 1/4 or 3/4, or by kind (= rarely, !=
    almost always, ordered ones at 1/4, 1/2 or 3/4), which bracket what
    the compares the histograms cannot estimate may do;
  - program shape: compares the histograms estimate and those they cannot,
    keep and drop branches, Invalid and logic ops, program length;
  - the tablet: rows, and per column its min, max and a 16-bin histogram
    over the table's range.
The network (include/virg_gpu.h, vg_mlp_*) is trained on the GPU by the
project's own kernels on per-tablet counts the executor reports
(vg_query_tablet_rows).  No counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import NP_TYPE, OP_Table, VIRG_SUCCESS, VirginianError, lib, make_ops

NCOLS = 6
NBINS = 16
NFEATURES = 4 + 7 + 1 + 4 + NCOLS * (2 + NBINS)
# feature 0 is logit(estimate) / SKIP; the network adds SKIP * feature 0 to
# its output's pre-activation and so learns a correction to the estimate
SKIP = 4.0


# ------------------------------------------------------------ tablet stats
def _column(meta, c: int) -> np.ndarray:
    dt = np.dtype(NP_TYPE[meta.fixed_type[c]])
    base = C.addressof(meta) + meta.fixed_block + meta.fixed_offset[c]
    buf = (C.c_char * (meta.rows * dt.itemsize)).from_address(base)
    return np.frombuffer(buf, dtype=dt)


class TableStats:
    """Per-tablet min, max and histogram of each fixed column, over the
    table's own range (computed once when the table is loaded)."""

    def __init__(self, db, table_id: int):
        tabs = db.tablets(table_id)
        self.rows = np.array([t.contents.rows for t in tabs], dtype=np.float64)
        ncol = min(NCOLS, tabs[0].contents.fixed_columns)
        self.is_int = [np.issubdtype(NP_TYPE[tabs[0].contents.fixed_type[c]], np.integer) for c in range(ncol)]
        mins = np.full((len(tabs), NCOLS), 0.0)
        maxs = np.full((len(tabs), NCOLS), 0.0)
        for i, t in enumerate(tabs):
            for c in range(ncol):
                col = _column(t.contents, c)
                if len(col):
                    mins[i, c], maxs[i, c] = float(col.min()), float(col.max())
        self.lo = mins.min(axis=0)
        self.hi = np.maximum(maxs.max(axis=0), self.lo + 1e-6)
        self.edges = [np.linspace(self.lo[c], self.hi[c], NBINS + 1) for c in range(NCOLS)]
        self.hist = np.zeros((len(tabs), NCOLS, NBINS))
        for i, t in enumerate(tabs):
            for c in range(ncol):
                col = _column(t.contents, c)
                if len(col):
                    h, _ = np.histogram(col, bins=self.edges[c])
                    self.hist[i, c] = h / len(col)
        span = self.hi - self.lo
        self.tablet_features = np.concatenate([
            ((mins - self.lo) / span)[:, :, None], ((maxs - self.lo) / span)[:, :, None], self.hist],
            axis=2).reshape(len(tabs), -1)

    def c_stats(self) -> "VgTableStats":
        """The statistics as vg_program_features takes them (arrays kept
        alive by this object)."""
        if getattr(self, "_c", None) is None:
            self._arrs = [np.ascontiguousarray(self.rows, dtype=np.float64),
                          np.ascontiguousarray(np.stack(self.edges), dtype=np.float64),
                          np.ascontiguousarray(self.hist, dtype=np.float64),
                          np.ascontiguousarray([int(x) for x in self.is_int] + [0] * (NCOLS - len(self.is_int)),
                                               dtype=np.int32),
                          np.ascontiguousarray(self.tablet_features, dtype=np.float64)]
            d = C.POINTER(C.c_double)
            r, e, h, i, t = self._arrs
            self._c = VgTableStats(len(self.rows), r.ctypes.data_as(d), e.ctypes.data_as(d), h.ctypes.data_as(d),
                                   i.ctypes.data_as(C.POINTER(C.c_int)), t.ctypes.data_as(d))
        return self._c


class VgTableStats(C.Structure):
    """vg_table_stats (include/virg_gpu.h)."""
    _fields_ = [("ntab", C.c_uint), ("rows", C.POINTER(C.c_double)), ("edges", C.POINTER(C.c_double)),
                ("hist", C.POINTER(C.c_double)), ("is_int", C.POINTER(C.c_int)),
                ("tabfeat", C.POINTER(C.c_double))]


# ------------------------------------------------------- program features
def program_features_ops(arr, n: int, stats: TableStats) -> np.ndarray:
    """Feature rows (one per tablet) of an op array (virg_op * n) over the
    table `stats` describes: vg_program_features in the host library."""
    _bind()
    out = np.empty((len(stats.rows), NFEATURES), dtype=np.float32)
    if lib.vg_program_features(arr, n, C.byref(stats.c_stats()),
                               out.ctypes.data_as(C.POINTER(C.c_float))) != VIRG_SUCCESS:
        raise VirginianError("vg_program_features failed")
    return out


def program_features(ops: list[tuple], stats: TableStats) -> np.ndarray:
    """Feature rows (one per tablet) of a program over the table `stats`
    describes (see include/virg_gpu.h, vg_program_features)."""
    arr, n, _keep = make_ops(ops)
    return program_features_ops(arr, n, stats)


# ------------------------------------------------------------------- MLP
class MLP:
    """The predictor network on the GPU (vg_mlp_* in include/virg_gpu.h)."""

    def __init__(self, nin: int = NFEATURES, nh: int = 32, seed: int = 1, skip: float = SKIP,
                 loss: str = "absolute", tau: float = 0.9):
        """loss: "absolute" (mean absolute error, what an estimate is
        judged by; the squared error pulls predictions toward the mean of
        skewed selectivities), "squared", or "quantile" (the pinball loss
        of quantile tau: the network learns an upper bound whose width
        follows how uncertain each program is, which is what sizing
        buffers for a target rerun rate needs)."""
        _bind()
        self.nin, self.nh, self.skip = nin, nh, skip
        self.absolute = loss == "absolute"
        self.tau = tau if loss == "quantile" else None
        self.h = lib.vg_mlp_create(nin, nh, skip, seed)
        if not self.h:
            raise VirginianError("vg_mlp_create failed (no device?)")
        ok = (lib.vg_mlp_set_quantile(self.h, tau) if self.tau is not None
              else lib.vg_mlp_set_loss(self.h, 1 if self.absolute else 0))
        if ok != VIRG_SUCCESS:
            raise VirginianError("setting the network's loss failed")
        self.nparams = lib.vg_mlp_num_params(self.h)

    def criterion(self, y: np.ndarray, t: np.ndarray) -> float:
        """The validation criterion: mean absolute error, or for a quantile
        network its pinball loss."""
        if self.tau is None:
            return float(np.mean(np.abs(y - t)))
        d = y.astype(np.float64) - t
        return float(np.mean(np.where(d > 0, (1.0 - self.tau) * d, -self.tau * d)))

    def params(self) -> np.ndarray:
        p = np.empty(self.nparams, dtype=np.float32)
        lib.vg_mlp_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if lib.vg_mlp_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))) != VIRG_SUCCESS:
            raise VirginianError("vg_mlp_set_params failed")

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(len(x), dtype=np.float32)
        if len(x) and lib.vg_mlp_forward(self.h, x.ctypes.data_as(C.POINTER(C.c_float)), len(x),
                                         y.ctypes.data_as(C.POINTER(C.c_float))) != VIRG_SUCCESS:
            raise VirginianError("vg_mlp_forward failed")
        return y

    def forward_host(self, x: np.ndarray) -> np.ndarray:
        """The same network on the host (vg_mlp_forward_host): no device
        call, so it does not wait for kernels that hold the compute units."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(len(x), dtype=np.float32)
        if len(x) and lib.vg_mlp_forward_host(self.h, x.ctypes.data_as(C.POINTER(C.c_float)), len(x),
                                              y.ctypes.data_as(C.POINTER(C.c_float))) != VIRG_SUCCESS:
            raise VirginianError("vg_mlp_forward_host failed")
        return y

    def gradient(self, x: np.ndarray, t: np.ndarray) -> tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = C.c_float(0)
        if lib.vg_mlp_gradient(self.h, x.ctypes.data_as(C.POINTER(C.c_float)),
                               t.ctypes.data_as(C.POINTER(C.c_float)), len(x), C.byref(loss),
                               g.ctypes.data_as(C.POINTER(C.c_float))) != VIRG_SUCCESS:
            raise VirginianError("vg_mlp_gradient failed")
        return loss.value, g

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int = 0, batch: int = 256, lr: float = 2e-3,
              seed: int = 1, steps: int = 600) -> float:
        """Adam on minibatches; epochs = 0: as many passes as `steps`
        updates take (the example sets are small: a few thousand tablets
        of a few hundred programs)."""
        if epochs <= 0:
            epochs = max(1, math.ceil(steps / max(1, math.ceil(len(x) / batch))))
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float(0)
        if lib.vg_mlp_train(self.h, x.ctypes.data_as(C.POINTER(C.c_float)),
                            t.ctypes.data_as(C.POINTER(C.c_float)), len(x), epochs, batch, lr, seed,
                            C.byref(loss)) != VIRG_SUCCESS:
            raise VirginianError("vg_mlp_train failed")
        return loss.value

    def fit(self, x: np.ndarray, t: np.ndarray, xv: np.ndarray, tv: np.ndarray, max_epochs: int = 300,
            patience: int = 60, batch: int = 256, lr: float = 2e-3, seed: int = 1) -> dict:
        """Train epoch by epoch (vg_mlp_train, Adam state carried over) and
        keep the parameters with the least criterion (mean absolute error,
        or the pinball loss of a quantile network) on the validation
        examples (xv, tv: tablets of programs kept out of x);
        stop after `patience` epochs without improvement (the first tens
        of epochs, while the correction grows from zero, may not improve
        it).  A few hundred programs are easy to memorise: without this
        the held-out error rises above the estimate the network corrects."""
        best = self.criterion(self.forward_host(xv), tv)
        start = best
        best_p, best_epoch, epoch = self.params(), 0, 0
        for epoch in range(1, max_epochs + 1):
            self.train(x, t, epochs=1, batch=batch, lr=lr, seed=seed + epoch)
            mae = self.criterion(self.forward_host(xv), tv)
            if mae < best:
                best, best_p, best_epoch = mae, self.params(), epoch
            elif epoch - best_epoch >= patience:
                break
        self.set_params(best_p)
        return {"validation_mae": best, "validation_mae_start": start, "best_epoch": best_epoch,
                "epochs": epoch}

    def close(self) -> None:
        if self.h:
            lib.vg_mlp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_BOUND = False


def _bind() -> None:
    global _BOUND
    if _BOUND:
        return
    F = C.POINTER(C.c_float)
    P = C.c_void_p
    sig = {
        "vg_mlp_create": (P, [C.c_uint, C.c_uint, C.c_float, C.c_ulonglong]),
        "vg_mlp_free": (None, [P]),
        "vg_mlp_num_params": (C.c_uint, [P]),
        "vg_mlp_get_params": (C.c_int, [P, F]),
        "vg_mlp_set_params": (C.c_int, [P, F]),
        "vg_mlp_forward": (C.c_int, [P, F, C.c_uint, F]),
        "vg_mlp_forward_host": (C.c_int, [P, F, C.c_uint, F]),
        "vg_mlp_gradient": (C.c_int, [P, F, F, C.c_uint, C.POINTER(C.c_float), F]),
        "vg_mlp_set_loss": (C.c_int, [P, C.c_int]),
        "vg_mlp_set_quantile": (C.c_int, [P, C.c_float]),
        "vg_mlp_train": (C.c_int, [P, F, F, C.c_uint, C.c_uint, C.c_uint, C.c_float, C.c_ulonglong,
                                   C.POINTER(C.c_float)]),
        "vg_query_tablet_rows": (C.c_uint, [P, C.POINTER(C.c_ulonglong), C.c_uint]),
        "vg_program_features": (C.c_int, [P, C.c_uint, P, F]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _BOUND = True


# ----------------------------------------------------- training and sizing
def table_of(ops: list[tuple]) -> int:
    return next(o[1] for o in ops if o[0] == OP_Table)


def tablet_rows_out(q) -> np.ndarray:
    """Rows q's last two-pass launch selected in each of its tablets."""
    _bind()
    n = 1 << 16
    buf = (C.c_ulonglong * n)()
    k = lib.vg_query_tablet_rows(q.h, buf, n)
    return np.array(buf[:k], dtype=np.float64)


def collect(db, programs: list[list[tuple]], stats: dict) -> tuple[np.ndarray, np.ndarray]:
    """Training examples from the executor's own counts: run each program
    in two passes and pair each tablet's features with the fraction of
    its rows selected."""
    xs, ts = [], []
    for ops in programs:
        st = stats[table_of(ops)]
        q = db.prepare(ops, flags=1)
        try:
            q.run()
            sel = tablet_rows_out(q) / np.maximum(st.rows, 1)
        finally:
            q.close()
        xs.append(program_features(ops, st))
        ts.append(sel)
    return np.concatenate(xs).astype(np.float32), np.concatenate(ts).astype(np.float32)


def capacity(pred: np.ndarray, rows: np.ndarray, margin: float) -> int:
    """Output rows to reserve: per tablet rows * min(1, pred + margin),
    plus a little slack."""
    return int(np.ceil(rows * np.minimum(1.0, np.maximum(pred, 0.0) + margin)).sum()) + 64


def calibrate(pred_sel: list[np.ndarray], true_rows: list[float], rows: list[np.ndarray],
              max_overflow: float) -> float:
    """Smallest margin (on a grid) with which at most max_overflow of the
    queries would not fit."""
    for m in np.linspace(0.0, 1.0, 201):
        over = sum(capacity(p, r, m) < t for p, t, r in zip(pred_sel, true_rows, rows))
        if over <= max_overflow * len(true_rows):
            return float(m)
    return 1.0
