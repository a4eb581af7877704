import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "tests", ROOT / "tests" / "golden", ROOT / "aero-cfd-solver_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP backend")


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(GOLDEN / f"golden_{name}.npz", allow_pickle=False)
    return load


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle_lib as ol
    ol.oracle()
    return ol
