"""The decoder's register-slot allocation, checked on the CPU over seeded
random programs.

The decoder gives a computed value a slot whose previous value has had
its last reader (loaded columns keep slots of their own).  The same
decoder built with VG_DECODE_PRIVATE_SLOTS gives every value a slot of its
own, so there each read finds exactly the value the program means.  For
every pass as the kernels run it (the count pass: filter columns, then the
filter's micro-ops; the write pass: the columns it reloads, the micro-ops it
needs, the next tile's columns when no output reads their slots, then the
outputs; the single pass: filter columns, filter micro-ops, projection
columns, the rest, the outputs), each read is traced to the load or
micro-op that last wrote its slot.  The two layouts must trace every read
to the same writer: a reused slot overwritten while its value is still
live traces to the wrong one.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

import virginian_amd as va
from virginian_amd import workload as W

ROOT = Path(__file__).resolve().parents[1]

UOP = np.dtype([("kind", "u1"), ("type", "u1"), ("flags", "u1"), ("target", "u1"), ("dst", "u1"),
                ("dst_hi", "u1"), ("a", "u1"), ("a_hi", "u1"), ("b", "u1"), ("b_hi", "u1"),
                ("src_op", "<u2"), ("wslot", "u1"), ("join", "u1"), ("pad", "<u2"),
                ("ia", "<u8"), ("ib", "<u8")])
COL = np.dtype([("src", "u1"), ("width", "u1"), ("phase", "u1"), ("type", "u1"), ("slot", "u1"),
                ("slot_hi", "u1"), ("wload", "u1"), ("pad", "u1")])
OUT = np.dtype([("kind", "u1"), ("width", "u1"), ("slot", "u1"), ("slot_hi", "u1"), ("col", "u1"),
                ("type", "u1"), ("pad", "<u2"), ("imm", "<u8")])
assert UOP.itemsize == 32 and COL.itemsize == 8 and OUT.itemsize == 16
# vg_prog: 8 u32 counters, wneed (u64), nslot_count, wsafe, then the arrays
HDR, NCOLS, NOUTS, NUOPS = 48, 18, 16, 48
K_NOT, K_INVALID, K_MOV, K_CAST = 8, 9, 14, 15
F_AIMM, F_BIMM = 1, 2
PHASE_FILTER = 0
OUT_SLOT = 0


@pytest.fixture(scope="module")
def private_decoder(tmp_path_factory):
    """decode.c built with every value in a slot of its own (links the
    product library only for the op names its error messages print)."""
    out = tmp_path_factory.mktemp("dec") / "libdecode_private.so"
    gpu = ROOT / "virginian_amd" / "csrc" / "gpu"
    product = Path(va.lib._name)
    subprocess.run(["gcc", "-O1", "-shared", "-fPIC", "-std=gnu11", "-DVG_DECODE_PRIVATE_SLOTS",
                    f"-I{ROOT / 'include'}", f"-I{ROOT / 'virginian_amd' / 'csrc' / 'host'}", f"-I{gpu}",
                    str(gpu / "decode.c"), "-o", str(out), str(product), f"-Wl,-rpath,{product.parent}"],
                   check=True)
    lib = C.CDLL(str(out))
    f = lib.vg_decode     # decode.c's entry; vg_decode_program wraps it in the product
    f.restype = C.c_int
    f.argtypes = [C.POINTER(va.Op), C.c_uint, C.c_uint, C.POINTER(va.TabletMeta), C.c_void_p, C.c_void_p,
                  C.c_char_p, C.c_size_t]
    addr = C.cast(f, C.c_void_p).value
    assert addr != C.cast(va.lib.vg_decode, C.c_void_p).value, "private build resolved to the product's"
    return lambda ops, n, pc, tab, out, size, err, errlen: f(ops, n, pc, tab, None, out, err, errlen)


def decode(fn, ops, tab):
    arr, n, keep = va.make_ops(ops)
    par = [i for i, o in enumerate(ops) if o[0] == va.OP_Parallel][0]
    buf = (C.c_char * 8192)()
    err = C.create_string_buffer(256)
    if fn(arr, n, par + 1, tab, buf, 8192, err, 256) != 1:
        return None, err.value.decode()
    raw = bytes(buf)
    h = np.frombuffer(raw[:32], dtype="<u4")
    nuop, nfilter, ncol, nload, nout, nslot, _, wfull = (int(x) for x in h)
    wneed = int(np.frombuffer(raw[32:40], dtype="<u8")[0])
    nslot_count, wsafe = (int(x) for x in np.frombuffer(raw[40:48], dtype="<u4"))
    o = HDR
    cols = np.frombuffer(raw[o:o + NCOLS * 8], dtype=COL)[:ncol]
    o += NCOLS * 8
    outs = np.frombuffer(raw[o:o + NOUTS * 16], dtype=OUT)[:nout]
    o += NOUTS * 16
    uops = np.frombuffer(raw[o:o + NUOPS * 32], dtype=UOP)[:nuop]
    return dict(nuop=nuop, nfilter=nfilter, nload=nload, nslot=nslot, nslot_count=nslot_count,
                wfull=wfull, wneed=wneed, wsafe=wsafe, cols=cols, outs=outs, uops=uops), None


def reads(u) -> list[int]:
    """Slots a micro-op reads (vm_kernel.hip: cond_strips, math_strips,
    cast_from; NOT reads only a, INVALID nothing, immediates no slot)."""
    kind = int(u["kind"])
    if kind == K_INVALID:
        return []
    r = [] if u["flags"] & F_AIMM else [int(u["a"]), int(u["a_hi"])]
    if kind not in (K_NOT, K_MOV, K_CAST) and not (u["flags"] & F_BIMM):
        r += [int(u["b"]), int(u["b_hi"])]
    return r


def trace(P, passname: str) -> list:
    """(reader, slot's writer) for every read of one pass, in order."""
    cols, uops, outs = P["cols"], P["uops"], P["outs"]
    nload, nu, nf = P["nload"], P["nuop"], P["nfilter"]
    owner: dict[int, tuple] = {}
    seen: list = []

    def load(c, tag="col"):
        owner[int(cols[c]["slot"])] = (tag, c)
        owner[int(cols[c]["slot_hi"])] = (tag, c)

    def run(k):
        u = uops[k]
        for j, s in enumerate(reads(u)):
            seen.append((("uop", k, j), owner.get(s)))
        if int(u["kind"]) > K_INVALID:
            owner[int(u["dst"])] = ("uop", k)
            owner[int(u["dst_hi"])] = ("uop", k)

    def emit():
        for j, o in enumerate(outs):
            if int(o["kind"]) == OUT_SLOT:
                seen.append((("out", j, 0), owner.get(int(o["slot"]))))
                if o["width"] == 8:
                    seen.append((("out", j, 1), owner.get(int(o["slot_hi"]))))

    filt = [c for c in range(nload) if cols[c]["phase"] == PHASE_FILTER]
    proj = [c for c in range(nload) if cols[c]["phase"] != PHASE_FILTER]
    if passname == "count":
        for c in filt:
            load(c)
        for k in range(nf):
            run(k)
    elif passname == "write":
        wl = [c for c in range(nload) if cols[c]["wload"]]
        for c in wl:
            load(c)
        for k in range(nu):
            if P["wfull"] or (P["wneed"] >> k) & 1:
                run(k)
        if P["wsafe"]:
            for c in wl:
                load(c, "next tile col")
        emit()
    else:
        for c in filt:
            load(c)
        for k in range(nf):
            run(k)
        for c in proj:
            load(c)
        for k in range(nf, nu):
            run(k)
        emit()
    return seen


def check_same_reads(ops, tab, private_decoder, where):
    """Slots the reusing layout saves (None: too many slots without reuse)."""
    mine, err = decode(va.lib.vg_decode_program, ops, tab)
    priv, perr = decode(private_decoder, ops, tab)
    assert (mine is None) == (err is not None)
    if priv is None:
        return None
    assert mine is not None, (where, err)
    assert mine["nslot"] <= priv["nslot"] and mine["nslot_count"] <= priv["nslot_count"]
    assert mine["wneed"] == priv["wneed"] and mine["wfull"] == priv["wfull"]
    for passname in ("count", "write", "single"):
        want = trace(priv, passname)
        got = trace(mine, passname)
        assert [w for w, _ in got] == [w for w, _ in want]
        for (at, a), (_, b) in zip(got, want):
            assert b is not None, (where, passname, at, "private layout reads an unwritten slot")
            assert a == b, (where, passname, at, a, b, ops)
        if passname == "write" and not mine["wfull"]:
            # the write pass loads exactly the columns whose values it reads
            read = {w[1] for _, w in got if w[0] == "col"}
            wl = {c for c in range(mine["nload"]) if mine["cols"][c]["wload"]}
            assert read == wl, (where, read, wl)
    return priv["nslot"] - mine["nslot"]


SQL = [
    "select (uniformf * 2.0) + normalf5 as a, uniformi - normali20 as b, "
    "uniformf / (normalf20 + 1000.0) as c from test where uniformf < 50.0",
    "select uniformi, normali5 * normali20 - uniformi as d from test where normali5 > 0 and uniformi < 40",
    "select uniformf + normalf5 + normalf20 as s, uniformf * uniformf as q, normali20 from test "
    "where uniformf > 10.0 or normalf5 < 0.0",
]


def test_slot_reuse_on_filter_math_queries(db, private_decoder):
    tid = db.gen_table("test", 5000, 1 << 20, seed=1)
    tab = db.tablets(tid)[0]
    saved = [check_same_reads(va.program(db.compile(q)), tab, private_decoder, q) for q in SQL]
    assert None not in saved and saved[0] > 0


def test_slot_reuse_never_clobbers_a_live_value(db, private_decoder):
    tid = db.gen_table("test", 5000, 1 << 20, seed=1)
    tab = db.tablets(tid)[0]
    rng = np.random.default_rng(123)
    compared = saved = 0
    for i in range(600):
        sel = float(rng.uniform(0.0001, 0.8)) if rng.random() < 0.8 else None
        ops = W.random_program(rng, max_ops=40, selectivity=sel, table=tid)
        got = check_same_reads(ops, tab, private_decoder, i)
        if got is None:
            continue            # too many slots without reuse; nothing to compare
        compared += 1
        saved += got
    assert compared >= 500
    assert saved > 0                # reuse does take effect on these programs
