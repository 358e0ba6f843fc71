"""ORACLE -- test infrastructure only (tests/ import it; the product never
does).  A numpy restatement of the result-size predictor's features,
vg_program_features in virginian_amd/csrc/host/features.c, which the tests
check the C code against.

The features have no counterpart in the reference; the control flow the
walk follows is the VM's (reference: src/vm/virginia.c:48-130, :424-440):
a taken compare jumps to p3, keeping validity when p4 is set and dropping
the row otherwise; a row that does not take it falls through; Invalid
drops the rows reaching it; jumps are forward only.  Each `column op
constant` compare passes at the column histogram's estimate on the tablet
(linear inside a bin; integer columns by their integer steps), any other
compare at a rate given by its kind (ordered, =, !=), in seven variants
(VARIANTS).

Pinned by: closed forms on uniform columns (uniformi < 50 keeps half of
each tablet, uniformf > 90 a tenth), a two-compare AND whose estimate is
the product of its parts, and the shape counts of the published queries
(tests/test_mlp.py).
"""
from __future__ import annotations

import math

import numpy as np

import virginian_amd as va

CMP_OPS = [va.OP_Lt, va.OP_Le, va.OP_Gt, va.OP_Ge, va.OP_Eq, va.OP_Neq]
MIRROR = {va.OP_Lt: va.OP_Gt, va.OP_Le: va.OP_Ge, va.OP_Gt: va.OP_Lt, va.OP_Ge: va.OP_Le,
          va.OP_Eq: va.OP_Eq, va.OP_Neq: va.OP_Neq}
NCOLS, NBINS, SKIP = 6, 16, 4.0
# passing rate of an unknown (ordered or logic, =, !=) compare, per variant
VARIANTS = [(0.5, 0.5, 0.5), (0.25, 0.25, 0.25), (0.75, 0.75, 0.75), (0.5, 0.05, 0.95), (0.25, 0.5, 0.5),
            (0.75, 0.5, 0.5), (0.5, 0.0, 1.0)]


def cdf(st, c: int, v: float) -> np.ndarray:
    """Fraction of each tablet's rows with column c < v."""
    e = st.edges[c]
    if v <= e[0]:
        return np.zeros(len(st.rows))
    if v >= e[-1]:
        return np.ones(len(st.rows))
    k = min(int(np.searchsorted(e, v, side="right")) - 1, NBINS - 1)
    part = (v - e[k]) / (e[k + 1] - e[k])
    return st.hist[:, c, :k].sum(axis=1) + part * st.hist[:, c, k]


def passing(st, c: int, op: int, v: float) -> np.ndarray:
    """Estimated fraction of each tablet's rows with (column c op v)."""
    if st.is_int[c]:
        lt, le = cdf(st, c, math.ceil(v)), cdf(st, c, math.floor(v) + 1)
        eq = le - cdf(st, c, math.floor(v)) if float(v).is_integer() else np.zeros(len(st.rows))
    else:
        lt = le = cdf(st, c, v)
        eq = np.zeros(len(st.rows))
    return {va.OP_Lt: lt, va.OP_Le: le, va.OP_Gt: 1 - le, va.OP_Ge: 1 - lt, va.OP_Eq: eq, va.OP_Neq: 1 - eq}[op]


def _value(op) -> float:
    return float(np.float32(op[4])) if op[0] == va.OP_Float else float(op[2])


def walk(ops: list[tuple], st, unknown):
    """Per variant of `unknown` (rows of (ordered, =, !=) passing rates),
    the fraction of each tablet's rows reaching Result valid; and the
    program's shape counts."""
    ntab = len(st.rows)
    rates = np.asarray(unknown, dtype=np.float64)
    u = rates[:, :1] * np.ones((1, ntab))
    start = next((i + 1 for i, o in enumerate(ops) if o[0] == va.OP_Parallel), 0)
    res = next((i for i, o in enumerate(ops) if o[0] == va.OP_Result and i >= start), len(ops))
    arrive = {start: np.ones_like(u)}
    regdef: dict[int, tuple] = {}
    known = nunknown = keep = drop = ninvalid = nlogic = 0
    for pc in range(start, res):
        o = ops[pc]
        m = arrive.pop(pc, None)
        k = o[0]
        nxt = m
        if k == va.OP_Column:
            regdef[o[1]] = ("col", o[2])
        elif k in (va.OP_Integer, va.OP_Float):
            regdef[o[1]] = ("const", _value(o))
        elif k == va.OP_Invalid:
            ninvalid += 1
            nxt = None
        elif k in CMP_OPS or k in (va.OP_And, va.OP_Or, va.OP_Not):
            kind = 1 if k == va.OP_Eq else 2 if k == va.OP_Neq else 0
            p = rates[:, kind:kind + 1] * np.ones((1, ntab))
            if k in CMP_OPS:
                a, b = regdef.get(o[1], ("expr",)), regdef.get(o[2], ("expr",))
                op = k
                if a[0] == "const" and b[0] == "col":
                    a, b, op = b, a, MIRROR[k]
                if a[0] == "col" and a[1] < NCOLS and b[0] == "const":
                    p = np.clip(passing(st, a[1], op, b[1]), 0.0, 1.0)[None, :]
                    known += 1
                else:
                    nunknown += 1
                if o[4] or 0:
                    keep += 1
                else:
                    drop += 1
            else:
                nlogic += 1
            if m is not None:
                if (o[4] or 0) and o[3] > pc:
                    arrive[o[3]] = arrive.get(o[3], 0) + m * p
                nxt = m * (1 - p)
        elif len(o) > 1 and k != va.OP_Table:
            regdef[o[1]] = ("expr",)
        if nxt is not None:
            arrive[pc + 1] = arrive.get(pc + 1, 0) + nxt
    est = arrive.get(res)
    est = np.zeros_like(u) if est is None else np.clip(est * np.ones_like(u), 0.0, 1.0)
    return est, (known, nunknown, keep, drop, ninvalid, nlogic)


def _logit(p: np.ndarray) -> np.ndarray:
    q = np.clip(p, 1e-4, 1 - 1e-4)
    return np.log(q / (1 - q)) / SKIP


def program_features(ops: list[tuple], st) -> np.ndarray:
    """Feature rows (one per tablet), the layout vg_program_features writes."""
    ntab = len(st.rows)
    ests, (known, unknown, keep, drop, ninvalid, nlogic) = walk(ops, st, VARIANTS)
    est, lo, hi = ests[0], ests[1], ests[2]
    shape = np.array([known / 4, unknown / 4, keep / 4, drop / 4, ninvalid, nlogic, len(ops) / 40.0])
    return np.concatenate([_logit(est)[:, None], _logit(lo)[:, None], _logit(hi)[:, None], est[:, None],
                           np.tile(shape, (ntab, 1)), (st.rows / float(1 << 20))[:, None],
                           np.stack([_logit(ests[v]) for v in range(3, len(VARIANTS))], axis=1),
                           st.tablet_features], axis=1).astype(np.float32)
