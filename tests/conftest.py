import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers",
                            "gpu: needs a HIP device (MI355X); runs the "
                            "kernels through the C ABI")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden():
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = load_golden(name)
        return cache[name]
    return get


def vec_rel_err(got, ref):
    """Max over particles of |got_i - ref_i| / |ref_i| for a tuple of
    component arrays, normalising each particle by the norm of its vector
    (the whole tuple), so near-zero components do not inflate the error."""
    g = np.stack([np.asarray(x, dtype=np.float64) for x in got])
    r = np.stack([np.asarray(x, dtype=np.float64) for x in ref])
    norm = np.sqrt((r * r).sum(axis=0))
    diff = np.sqrt(((g - r) ** 2).sum(axis=0))
    scale = np.where(norm > 0, norm, 1.0)
    return float(np.max(diff / scale)) if g.size else 0.0


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
