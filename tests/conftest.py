import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (only to ask the runtime; nothing is computed)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def wc():
    """The product library; a GPU test fails loudly if it cannot load."""
    import libwetcloth_amd as W
    W.lib()
    return W


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den) if a.size else 0.0
