"""Build the native libraries in-tree.

    python -m virginian_amd.build            # product library + oracle
    python -m virginian_amd.build --product  # product library only

Product: virginian_amd/lib/libvirginian_amd.so -- host C library (gcc),
gfx950 kernels and the device layer (hipcc).  All code is compiled with
-ffp-contract=off so float arithmetic matches the reference VM's x86 SSE
arithmetic operation by operation.

Oracle (test infrastructure): oracle/liboracle.so from oracle/virg_oracle.c,
and, when the reference sources are present, oracle/_ref/ (see
oracle/build_ref.sh).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
OUT = PKG / "lib"
OBJ = PKG / "lib" / "obj"
LIBNAME = "libvirginian_amd.so"

HOST_C = ["db.c", "dbfile.c", "tablet.c", "table.c", "vm.c", "reader.c", "sql.c", "gen.c", "features.c"]
GPU_C = ["decode.c"]
GPU_HIP = ["vm_kernel.hip", "device.hip", "mlp.hip", "gen.hip", "jit.hip", "serve.hip"]

CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-Wall", "-Wextra",
          "-Wno-unused-parameter", "-pthread", f"-I{INCLUDE}", f"-I{CSRC / 'host'}",
          f"-I{CSRC / 'gpu'}"]
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            # no v_bitop3: this compiler has folded three-input mask logic into
            # it with a wrong truth table (the specialised kernels' hiprtc
            # options say the same)
            "-Xclang", "-target-feature", "-Xclang", "-bitop3-insts",
            "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC / 'host'}", f"-I{CSRC / 'gpu'}", f"-I{OBJ}"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the gfx950 kernels)")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "bitop3-insts" not in l
                    and "is not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out)


def embed_headers() -> None:
    """lib/obj/vgx_headers.inc: vg_prog.h and vgx_abi.h as one C string
    literal, the text jit.hip puts in front of every kernel it compiles at
    run time, so those kernels see the structs the host library uses."""
    text = "".join((CSRC / "gpu" / h).read_text() for h in ("vg_prog.h", "vgx_abi.h"))
    assert ")VGXH\"" not in text
    inc = OBJ / "vgx_headers.inc"
    body = 'R"VGXH(' + text + ')VGXH"\n'
    if not inc.exists() or inc.read_text() != body:
        inc.write_text(body)


def build_product(jobs: int = 8) -> Path:
    OUT.mkdir(parents=True, exist_ok=True)
    OBJ.mkdir(parents=True, exist_ok=True)
    embed_headers()
    hipcc = _hipcc()
    steps = []
    objs = []
    for f in HOST_C:
        o = OBJ / (Path(f).stem + ".o")
        steps.append(["gcc", *CFLAGS, "-c", str(CSRC / "host" / f), "-o", str(o)])
        objs.append(o)
    for f in GPU_C:
        o = OBJ / (Path(f).stem + ".o")
        steps.append(["gcc", *CFLAGS, "-c", str(CSRC / "gpu" / f), "-o", str(o)])
        objs.append(o)
    for f in GPU_HIP:
        o = OBJ / (Path(f).stem + ".hip.o")
        steps.append([hipcc, *HIPFLAGS, "-c", str(CSRC / "gpu" / f), "-o", str(o)])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for fut in [ex.submit(_run, s) for s in steps]:
            fut.result()
    lib = OUT / LIBNAME
    tmp = OUT / (LIBNAME + ".tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib", "-lpthread", "-lm"])
    os.replace(tmp, lib)
    return lib


def build_oracle() -> Path:
    src = ROOT / "oracle" / "virg_oracle.c"
    out = ROOT / "oracle" / "liboracle.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-ffp-contract=off", "-pthread",
          f"-I{INCLUDE}", str(src), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_reference_oracle() -> Path | None:
    """Compile the reference CPU VM from its own sources when they are
    present (this container only); the outputs go to oracle/_ref/."""
    script = ROOT / "oracle" / "build_ref.sh"
    if not Path("/root/reference/src/vm/virginia.c").exists() or not script.exists():
        return None
    _run(["bash", str(script)])
    lib = ROOT / "oracle" / "_ref" / "libvirgref.so"
    return lib if lib.exists() else None


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--product", action="store_true", help="product library only")
    ap.add_argument("-j", "--jobs", type=int, default=8)
    a = ap.parse_args(argv)
    print(build_product(a.jobs))
    if not a.product:
        print(build_oracle())
        ref = build_reference_oracle()
        if ref:
            print(ref)
    return 0


if __name__ == "__main__":
    sys.exit(main())
