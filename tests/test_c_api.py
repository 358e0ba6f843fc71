"""The public C API as a C program uses it: tests/c/api_example.c built
against include/ and linked with the product library (the condition
queries of src/test/sql.cc through virg_query / the reader)."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

import virginian_amd as va

ROOT = Path(__file__).resolve().parent.parent


def _build(tmp_path: Path) -> Path:
    exe = tmp_path / "api_example"
    lib = va._LIB_PATH
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "c" / "api_example.c"), "-o", str(exe),
                    "-L", str(lib.parent), "-lvirginian_amd", f"-Wl,-rpath,{lib.parent}"],
                   check=True)
    return exe


def test_c_program_builds_and_runs_without_gpu(tmp_path):
    dbfile = tmp_path / "test.db"
    out = subprocess.run([str(_build(tmp_path)), str(dbfile)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.split("\n")
    assert lines[0] == "rows 10"
    assert lines[1].startswith("ops ")
    # virg_close wrote the database: the meta block, the entry table, one tablet
    sizes = va.struct_sizes()           # virg_gpu_getsizes order: [6] virg_tablet_info, [11] virg_db
    assert dbfile.stat().st_size > sizes[11] + 16 * sizes[6]


@pytest.mark.gpu
def test_c_program_queries(tmp_path):
    out = subprocess.run([str(_build(tmp_path)), str(tmp_path / "test.db")], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    got = [int(l.split()[1]) for l in out.stdout.split("\n") if l.startswith("q")]
    assert got == [3, 2, 1, 1, 1, 1, 1, 3, 2]
