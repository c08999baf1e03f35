import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: full-size workloads")


@pytest.fixture(scope="session")
def ctx():
    """One solver context for the whole GPU session."""
    from kc_cuda_amd import Context
    c = Context(0)
    yield c
    c.close()


def canon_pairs(pairs: np.ndarray) -> np.ndarray:
    p = np.asarray(pairs, dtype=np.uint32).reshape(-1, 2)
    a = np.minimum(p[:, 0], p[:, 1])
    b = np.maximum(p[:, 0], p[:, 1])
    out = np.stack([a, b], axis=1)
    if out.shape[0] == 0:
        return out
    order = np.lexsort((out[:, 1], out[:, 0]))
    return out[order]
