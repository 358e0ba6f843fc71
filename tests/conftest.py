import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
# Prepared queries run on the two-pass interpreter unless a test asks for
# the specialised kernels (tests/test_gpu_jit.py sets the policy itself):
# compiling one per random program would dominate the suite's time.
os.environ.setdefault("VG_JIT", "0")
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_present() -> bool:
    """True when the library found a HIP device."""
    import virginian_amd as va
    db = va.Virginian()
    try:
        return db.device()[1] > 0
    finally:
        db.close()


@pytest.fixture
def db():
    import virginian_amd as va
    d = va.Virginian()
    yield d
    d.close()


@pytest.fixture
def gpu_db():
    import virginian_amd as va
    d = va.Virginian()
    if d.device()[1] == 0:
        d.close()
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X box")
    yield d
    d.close()
