import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # Autotuning results of this test session go to a private file.
    import tempfile
    os.environ.setdefault("T1M_TUNE_CACHE",
                          os.path.join(tempfile.mkdtemp(prefix="t1m_tune_"), "tuning.txt"))


@pytest.fixture(scope="session")
def kats():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")) as f:
        return json.load(f)
