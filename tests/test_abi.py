"""The C-ABI library on a host without a GPU: it loads, exports every
function the public headers declare, and its structs, opcodes and type
numbers are the reference's (tests/golden/abi.json, read from the
reference's own header by tests/golden/make_abi_fixtures.sh)."""
from __future__ import annotations

import ctypes as C
import json
import re
import subprocess
from pathlib import Path

import pytest

import virginian_amd as va

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = json.loads((ROOT / "tests" / "golden" / "abi.json").read_text())


def declared_functions() -> dict[str, list[str]]:
    """Function names declared in include/*.h (prototypes ending in ');')."""
    out = {}
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", " ", h.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        text = re.sub(r"#[^\n]*", " ", text)
        names = []
        for stmt in text.split(";"):
            s = " ".join(stmt.split())
            m = re.match(r"^(?:extern\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\((?:[^()]|\([^()]*\))*\)$",
                         s.split("{")[-1].strip())
            if m and not s.startswith("typedef") and "(*" not in s:
                names.append(m.group(1))
        out[h.name] = names
    return out


def test_headers_declare_boundary():
    decl = declared_functions()
    assert "virg_vm_gpu" in decl["virginian.h"]
    assert "virg_gpu_getsizes" in decl["virginian.h"]
    assert "vg_query_create" in decl["virg_gpu.h"]


def test_every_declared_symbol_is_exported():
    exported = set(subprocess.run(["nm", "-D", "--defined-only", str(va._LIB_PATH)],
                                  capture_output=True, text=True, check=True).stdout.split())
    missing = [(h, n) for h, names in declared_functions().items() for n in names
               if n not in exported]
    assert not missing, missing
    for names in declared_functions().values():
        for n in names:
            getattr(va.lib, n)          # resolvable through the dynamic loader


def test_struct_sizes_match_reference():
    """reference: src/test/virgstate.cc:64-80 compares these 15 sizes
    between the host code and the GPU code; here against the reference's
    own header."""
    assert va.struct_sizes() == GOLDEN["struct_sizes"]
    mirror = [C.sizeof(C.c_int), C.sizeof(C.c_float), C.sizeof(C.c_longlong),
              C.sizeof(C.c_double), C.sizeof(C.c_char), C.sizeof(va.TabletMeta),
              C.sizeof(va.TabletInfo), C.sizeof(va.Var), C.sizeof(va.VM), C.sizeof(va.Op),
              C.sizeof(va.ResultNode), C.sizeof(va.DB), C.sizeof(va.Reader)]
    want = GOLDEN["struct_sizes"]
    assert mirror == want[:13]
    assert C.sizeof(va.State) == want[14]


def test_numbering_matches_reference():
    for name, val in GOLDEN["opcodes"].items():
        assert getattr(va, name) == val, name
    for name, val in GOLDEN["types"].items():
        assert getattr(va, name) == val, name
    assert va.VIRG_SUCCESS == GOLDEN["status"]["VIRG_SUCCESS"]
    assert va.VIRG_FAIL == GOLDEN["status"]["VIRG_FAIL"]
    for k in ("VIRG_OPS", "VIRG_REGS", "VIRG_MAX_COLUMNS", "VIRG_TABLET_SIZE",
              "VIRG_MEM_TABLETS", "VIRG_ROW_BUFFER"):
        assert getattr(va, k) == GOLDEN["limits"][k], k
    for t, size in enumerate(GOLDEN["type_sizes"]):
        assert C.sizeof(C.c_int) and va.lib.virg_sizeof(t) == size


def test_header_compiles_as_c_with_reference_layout(tmp_path):
    """include/virginian.h is a plain C header: a C program built against it
    sees the reference's sizes and opcode numbers."""
    src = tmp_path / "probe.c"
    src.write_text(
        '#include <stdio.h>\n#include "virginian.h"\n#include "virg_gpu.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %d %d\\n\", sizeof(virg_tablet_meta),"
        " sizeof(virg_vm), sizeof(virg_op), sizeof(virg_db), sizeof(virginian), OP_Nop, VIRG_CHAR);"
        "return 0;}\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src),
                    "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    s = GOLDEN["struct_sizes"]
    assert got == [str(s[5]), str(s[8]), str(s[9]), str(s[11]), str(s[14]),
                   str(GOLDEN["opcodes"]["OP_Nop"]), str(GOLDEN["types"]["VIRG_CHAR"])]


def test_no_cpu_fallback_without_device():
    """Without a HIP device the executor refuses to run (there is no CPU
    path in the product)."""
    db = va.Virginian()
    try:
        if db.device()[1] > 0:
            pytest.skip("a GPU is present")
        db.gen_table("test", 1000, 0, seed=1)
        with pytest.raises(va.VirginianError):
            db.query("select id from test where uniformi < 50")
        with pytest.raises(va.VirginianError):
            db.prepare("select id from test where uniformi < 50")
    finally:
        db.close()
