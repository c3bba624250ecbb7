import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
HERE = Path(__file__).resolve().parent
if str(HERE) not in sys.path:
    sys.path.insert(0, str(HERE))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libddgx.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no HIP device is visible")
    return torch.device("cuda", 0)


def unit_square(Kx, Ky, N, jitter=0.0, seed=0):
    """Product-built mesh and the oracle mesh on identical vertices."""
    from ddg_amd import mesh as pm
    from oracle import mesh as om
    VX, VY, EToV = pm.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, Kx, Ky)
    if jitter:
        VX, VY = pm.jitter_vertices(VX, VY, 1.0 / max(Kx, Ky), jitter, seed)
    return pm.build_mesh_2d(VX, VY, EToV, N), om.build_mesh_2d(VX, VY, EToV, N)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
