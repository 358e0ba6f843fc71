"""Seeded synthetic workloads over the 7-column benchmark schema.

random_program() builds a virg_op program the way the reference's code
generator shapes them (src/sql/generate.c:591-776): Table, ResultColumn...,
Parallel, a filter section of forward-jumping compares, a projection
section, Result, Converge, Finish.  Programs are type-consistent (the
reference asserts equal operand types, src/vm/virginia.c:52, :143), never
divide integers and never make NaNs, and every result register is written
on every path that reaches Result, so any VM must give the same logical
result.  Filter compares either drop the row (jump to Result with p4 = 0)
or short-circuit an OR group (jump forward keeping validity, p4 = 1), and
the filter may redefine a temporary that the projection then reads, so
rows that skipped the redefinition keep the older value.
"""
from __future__ import annotations

import math

import numpy as np

from . import (OP_Add, OP_And, OP_Column, OP_Converge, OP_Div, OP_Eq, OP_Finish, OP_Float,
               OP_Ge, OP_Gt, OP_Integer, OP_Invalid, OP_Le, OP_Lt, OP_Mul, OP_Neq, OP_Not, OP_Or,
               OP_Parallel, OP_Result, OP_ResultColumn, OP_Rowid, OP_Sub, OP_Table, VIRG_FLOAT,
               VIRG_INT, VIRG_INT64, VIRG_OPS)

# benchmark schema: column index -> (name, type)
SCHEMA = [("uniformi", VIRG_INT), ("normali5", VIRG_INT), ("normali20", VIRG_INT),
          ("uniformf", VIRG_FLOAT), ("normalf5", VIRG_FLOAT), ("normalf20", VIRG_FLOAT)]
UNIFORMF = 3
CMPS = [OP_Lt, OP_Le, OP_Gt, OP_Ge, OP_Eq, OP_Neq]


class _Builder:
    def __init__(self, rng: np.random.Generator, nres: int):
        self.rng = rng
        self.ops: list[tuple] = []
        self.regtype: dict[int, int] = {}
        # result registers are the top nres registers; temporaries below,
        # handed out round-robin (an expression never needs more than the
        # ring holds) and never the pinned shared temporary
        self.res_regs = list(range(16 - nres, 16))
        self.ring = list(range(0, 16 - nres))
        self.pos = 0
        self.pinned: set[int] = set()

    def emit(self, op, p1=0, p2=0, p3=0, p4=None) -> int:
        self.ops.append((op, p1, p2, p3, p4))
        return len(self.ops) - 1

    def temp(self) -> int:
        while True:
            r = self.ring[self.pos % len(self.ring)]
            self.pos += 1
            if r not in self.pinned:
                return r

    def load(self, col: int, reg: int | None = None) -> int:
        r = self.temp() if reg is None else reg
        self.emit(OP_Column, r, col)
        self.regtype[r] = SCHEMA[col][1]
        return r

    def const(self, typ: int, reg: int | None = None) -> int:
        r = self.temp() if reg is None else reg
        if typ == VIRG_INT:
            self.emit(OP_Integer, r, int(self.rng.integers(-60, 100)))
        else:
            self.emit(OP_Float, r, 0, 0, float(np.float32(self.rng.uniform(-60.0, 100.0))))
        self.regtype[r] = typ
        return r

    def value(self, typ: int, depth: int, reg: int | None = None) -> int:
        """A register of type typ holding a column, a constant or an
        arithmetic combination of them."""
        rng = self.rng
        cols = [c for c, (_, t) in enumerate(SCHEMA) if t == typ]
        if depth <= 0 or rng.random() < 0.4:
            return self.load(int(rng.choice(cols)), reg)
        a = self.value(typ, depth - 1)
        kinds = [OP_Add, OP_Sub, OP_Mul] + ([OP_Div] if typ == VIRG_FLOAT else [])
        k = int(rng.choice(kinds))
        if k == OP_Div:
            # divide by a nonzero constant only: no NaN, no infinity
            b = self.temp()
            v = float(np.float32(rng.choice([-1.0, 1.0]) * rng.uniform(0.5, 50.0)))
            self.emit(OP_Float, b, 0, 0, v)
            self.regtype[b] = typ
        elif rng.random() < 0.5:
            b = self.const(typ)
        else:
            b = self.load(int(rng.choice(cols)))
        r = self.temp() if reg is None else reg
        self.emit(k, r, a, b)
        self.regtype[r] = typ
        return r


def random_program(rng: np.random.Generator, max_ops: int = VIRG_OPS, selectivity: float | None = None,
                   table: int = 0) -> list[tuple]:
    """One program as (op, p1, p2, p3, p4) tuples, at most max_ops ops."""
    for _ in range(200):
        prog = _try_program(rng, selectivity, table)
        if len(prog) <= max_ops:
            return prog
    raise RuntimeError("could not build a program within the op limit")


def _try_program(rng: np.random.Generator, selectivity: float | None, table: int) -> list[tuple]:
    nres = int(rng.integers(1, 4))
    b = _Builder(rng, nres)
    res_types = []
    for j in range(nres):
        res_types.append(int(rng.choice([VIRG_INT, VIRG_FLOAT, VIRG_INT64])) if j == 0
                         else int(rng.choice([VIRG_INT, VIRG_FLOAT])))
    b.emit(OP_Table, table)
    for j, t in enumerate(res_types):
        b.emit(OP_ResultColumn, t, 0, 0, f"r{j}")
    par = b.emit(OP_Parallel)
    fixups_result: list[int] = []       # compares jumping to Result
    fixups_proj: list[int] = []         # compares jumping to the projection start
    shared = None                       # a temporary the filter may redefine
    if selectivity is not None:
        u = b.load(UNIFORMF)
        c = b.temp()
        b.emit(OP_Float, c, 0, 0, float(np.float32(100.0 * selectivity)))
        b.regtype[c] = VIRG_FLOAT
        fixups_result.append(b.emit(OP_Ge, u, c, 0, 0))
    ngroups = int(rng.integers(0, 4))
    for g in range(ngroups):
        typ = int(rng.choice([VIRG_INT, VIRG_FLOAT]))
        if shared is None and rng.random() < 0.5:
            shared = (b.value(typ, 1), typ)
            b.pinned.add(shared[0])
        nterms = int(rng.integers(1, 4))
        group_fix: list[int] = []
        for t in range(nterms):
            a = b.value(typ, int(rng.integers(0, 2)))
            c = b.const(typ) if rng.random() < 0.7 else b.value(typ, 0)
            op = int(rng.choice(CMPS))
            if typ == VIRG_FLOAT and op in (OP_Eq, OP_Neq):
                op = OP_Lt
            if nterms == 1:
                # AND term: rows failing it are dropped (negated compare)
                fixups_result.append(b.emit(_negate(op), a, c, 0, 0))
            elif t < nterms - 1:
                # OR term: rows passing skip the rest of the group
                group_fix.append(b.emit(op, a, c, 0, 1))
            else:
                # last OR term: drop the rows that failed every term
                fixups_result.append(b.emit(_negate(op), a, c, 0, 0))
            if shared is not None and shared[1] == typ and rng.random() < 0.3:
                # redefinition: only rows still at this op update it
                k = b.const(typ)
                b.emit(OP_Add, shared[0], shared[0], k)
        end = len(b.ops)
        for i in group_fix:
            op, p1, p2, _, p4 = b.ops[i]
            b.ops[i] = (op, p1, p2, end, p4)
    if rng.random() < 0.25:
        # logic ops (never emitted by the reference's code generator, but
        # part of its VM): Not drops rows whose value is zero, And / Or
        # keep-branch to the next op
        typ = int(rng.choice([VIRG_INT, VIRG_FLOAT]))
        a = b.value(typ, int(rng.integers(0, 2)))
        c = b.value(typ, 0) if rng.random() < 0.5 else b.const(typ)
        k = int(rng.choice([OP_And, OP_Or, OP_Not]))
        if k == OP_Not:
            fixups_result.append(b.emit(OP_Not, a, 0, 0, 0))
        else:
            nxt = len(b.ops) + 1
            b.emit(k, a, c, nxt, 1)
    if rng.random() < 0.15:
        # the codegen's trailing pattern: a keep-branch over an Invalid
        typ = int(rng.choice([VIRG_INT, VIRG_FLOAT]))
        a = b.value(typ, 0)
        c = b.const(typ)
        i = b.emit(int(rng.choice([OP_Lt, OP_Gt, OP_Ge, OP_Le])), a, c, 0, 1)
        b.emit(OP_Invalid)
        fixups_proj.append(i)
    proj = len(b.ops)
    for i in fixups_proj:
        op, p1, p2, _, p4 = b.ops[i]
        b.ops[i] = (op, p1, p2, proj, p4)
    for j, t in enumerate(res_types):
        r = b.res_regs[j]
        if t == VIRG_INT64:
            b.emit(OP_Rowid, r)
            b.regtype[r] = VIRG_INT64
        elif shared is not None and shared[1] == t and rng.random() < 0.5:
            k = b.const(t)
            b.emit(OP_Add, r, shared[0], k)
            b.regtype[r] = t
        else:
            b.value(t, int(rng.integers(0, 3)), reg=r)
    res = b.emit(OP_Result, b.res_regs[0], nres)
    for i in fixups_result:
        op, p1, p2, _, p4 = b.ops[i]
        b.ops[i] = (op, p1, p2, res, p4)
    b.emit(OP_Converge)
    fin = b.emit(OP_Finish)
    op, p1, p2, _, p4 = b.ops[par]
    b.ops[par] = (op, p1, p2, fin, p4)
    return b.ops


def _negate(op: int) -> int:
    return {OP_Lt: OP_Ge, OP_Ge: OP_Lt, OP_Le: OP_Gt, OP_Gt: OP_Le, OP_Eq: OP_Neq, OP_Neq: OP_Eq}[op]


def log_uniform_sizes(rng: np.random.Generator, n: int, lo: float, hi: float) -> list[int]:
    return [int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(n)]


# The reference's published benchmark queries (src/test/queries.h:1-13),
# over the table suite_table() builds.
SUITE_QUERIES = [
    "select id, uniformi, normali5 from test where uniformi > 60 and normali5 < 0",
    "select id, uniformf, normalf5 from test where uniformf > 60.0 and normalf5 < 0.0",
    "select id, uniformi, normali5 from test where uniformi > -60 and normali5 < 5",
    "select id, uniformf, normalf5 from test where uniformf > -60.0 and normalf5 < 5.0",
    "select id, uniformi, normali20 from test where (normali20 + 40) > (uniformi - 10)",
    "select id, uniformf, normalf20 from test where (normalf20 + 40.0) > (uniformf - 10.0)",
    "select id, normali5, normali20 from test where normali5 * normali20 >= -5 and normali5 * normali20 <= 5",
    "select id, normalf5, normalf20 from test where normalf5 * normalf20 >= -5.0 and normalf5 * normalf20 <= 5.0",
    "select id, uniformi, normali5, normali20 from test where uniformi >= -1 and uniformi <= 1 or "
    "normali5 >= -1 and normali5 <= 1 or normali20 >= -1 and normali20 <= 1",
    "select id, uniformf, normalf5, normalf20 from test where uniformf >= -1.0 and uniformf <= 1.0 or "
    "normalf5 >= -1.0 and normalf5 <= 1.0 or normalf20 >= -1.0 and normalf20 <= 1.0",
]


def suite_table(db, rows: int, tablet_rows: int, seed: int, name: str = "test") -> int:
    """The reference's test table (db/generate.c:41-48, 8M rows in its
    benchmark, src/Makefile:85): int key 0..rows-1; uniformi, normali5,
    normali20 = C int casts (toward zero) of U(-100, 100), N(0, 5),
    N(0, 20); uniformf, normalf5, normalf20 = the same distributions as
    f32.  Seeded numpy draws stand in for the reference's gsl ranlxd2
    seeded by the clock.  Returns the table id."""
    rng = np.random.default_rng(seed)
    tid = db.create_table(name, VIRG_INT, SCHEMA)
    for r0 in range(0, rows, tablet_rows):
        n = min(tablet_rows, rows - r0)
        key = np.arange(r0, r0 + n, dtype=np.int32)
        cols = [np.trunc(rng.uniform(-100, 100, n)).astype(np.int32),
                np.trunc(rng.normal(0, 5, n)).astype(np.int32),
                np.trunc(rng.normal(0, 20, n)).astype(np.int32),
                rng.uniform(-100, 100, n).astype(np.float32),
                rng.normal(0, 5, n).astype(np.float32),
                rng.normal(0, 20, n).astype(np.float32)]
        db.append_tablet(tid, key, cols)
    return tid

