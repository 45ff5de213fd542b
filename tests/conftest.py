import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X)")


def golden(name):
    """Load a golden fixture (data only, no pickles)."""
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def graphene_sample(L, rescale=9.0, **kw):
    import tbplas_amd as tb
    s = tb.make_graphene_sample(L, **kw)
    s.rescale_ham(rescale)
    return s


def tau_of(sample, factor=1.0):
    return factor * 2 * math.pi / sample.energy_range * sample.rescale


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: GPU tests must see a GPU (no CPU fallback)."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch
