"""The committed workload fixtures (tests/golden/workload_digests.json) on
the CPU: the 1M-row entries are recomputed from the oracle here, and the
file records that the reference's own CPU VM agreed at 1M rows when it was
made; the CPU calibration file is complete."""
from __future__ import annotations

import json
from pathlib import Path

from oracle import oracle
from tests.helpers import CONFIG1_SQL, config2_sql
from virginian_amd.verify import Digest, digest

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_digest_streams():
    import numpy as np
    a, b = np.arange(10, dtype=np.int32), np.linspace(0, 1, 10, dtype=np.float32)
    d = Digest(2)
    d.update([a[:3], b[:3]])
    d.update([a[3:], b[3:]])
    assert d.hexdigest() == digest([a, b]) and d.rows == 10
    assert digest([a, b]) != digest([a, b[::-1].copy()])


def test_1m_fixtures_match_the_oracle(db):
    g = json.loads((GOLDEN / "workload_digests.json").read_text())
    assert g["seed"] == 1
    db.gen_table("test", 1 << 20, 1 << 20, seed=1, threads=8)
    got, _ = oracle.select(db, CONFIG1_SQL, threads=8)
    assert digest(list(got.values())) == g["config1"]["digest"]
    assert len(got["id"]) == g["config1"]["selected"] == sum(g["config1"]["per_tablet"])
    for q in g["config2_1M"]["queries"]:
        got, _ = oracle.select(db, config2_sql(q["threshold"]), threads=8)
        assert digest(list(got.values())) == q["digest"], q["threshold"]
    assert g["reference_check_1M"]["equal"] == [True] * 5
    for key in ("config2", "config3"):
        for q in g[key].get("queries", [g[key]]):
            assert q["selected"] == sum(q["per_tablet"])
    assert len(g["config3"]["per_tablet"]) == 1908


def test_cpu_calibration_present():
    c = json.loads((GOLDEN / "cpu_calibration.json").read_text())
    for name in ("config2_T10", "config1"):
        for mode in ("multi", "single"):
            r = c["queries"][name][mode]
            assert r["reference_ms"] > 0 and r["restatement_ms"] > 0
            assert abs(r["restatement_over_reference"] - r["restatement_ms"] / r["reference_ms"]) < 0.01
