"""Write tests/golden/ref_db_20000.db.gz: a database file written by the
reference's own code (oracle/_ref, built from /root/reference by
oracle/build_ref.sh -- in the development container only), holding the
7-column benchmark table "test" of 20,000 rows from the seeded generator
(seed 7).  tests/test_gpu_dbfile.py opens it on the GPU host, where the
reference never runs.

    python tests/golden/make_ref_db.py
"""
from __future__ import annotations

import gzip
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

ROWS, SEED = 20_000, 7
COLS = ("uniformi", "normali5", "normali20", "uniformf", "normalf5", "normalf20")
OUT = Path(__file__).resolve().parent / f"ref_db_{ROWS}.db.gz"


def main() -> int:
    from oracle import ref
    from tests.test_oracle import gen_columns
    if not ref.available():
        print("oracle/_ref/libvirgref.so is not built (oracle/build_ref.sh)", file=sys.stderr)
        return 1
    c = gen_columns(ROWS, SEED)
    with tempfile.TemporaryDirectory() as d:
        path = Path(d) / "ref.db"
        r = ref.RefDB(path=path)
        r.table("test", c["id"], [(n, c[n]) for n in COLS])
        r.write_close()
        OUT.write_bytes(gzip.compress(path.read_bytes(), 9, mtime=0))
    print(OUT, OUT.stat().st_size)
    return 0


if __name__ == "__main__":
    sys.exit(main())
