import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def torch_dev():
    """torch on the HIP device; the whole GPU tier requires it."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def scale_err(got, ref):
    """|s - 1| for the least-squares scale s = <ref, got> / <ref, ref>: a
    systematic scaling of the whole result shows in full, random rounding
    averages out over the cells."""
    g = np.asarray(got).ravel().astype(np.complex128)
    r = np.asarray(ref).ravel().astype(np.complex128)
    return float(abs(np.vdot(r, g) / np.vdot(r, r) - 1.0))
