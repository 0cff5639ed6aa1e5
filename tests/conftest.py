import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib

    oracle_lib.lib()
    return oracle_lib


@pytest.fixture(scope="session")
def steps():
    """The HIP library; fails loudly (never skips to a CPU path) when absent."""
    import steps_amd

    steps_amd.load_library()
    if steps_amd.device_count() < 1:
        pytest.fail("no HIP device visible: the gpu tests need an MI355X")
    return steps_amd


def rel_err(F, Fref):
    """Per-row relative force error |F - Fref| / |Fref|."""
    import numpy as np

    F = np.asarray(F, dtype=np.float64).reshape(-1, 3)
    Fr = np.asarray(Fref, dtype=np.float64).reshape(-1, 3)
    num = np.linalg.norm(F - Fr, axis=1)
    den = np.linalg.norm(Fr, axis=1)
    return num / np.where(den > 0, den, 1.0)


def report(tag, e):
    import numpy as np

    print(f"[{tag}] median {np.median(e):.3e} p99 {np.percentile(e, 99):.3e} max {np.max(e):.3e}", flush=True)
