import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libmfv.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.lib()
    return oracle


def make_tri_domain(nx, ny, lx=1.0, ly=1.0, jitter=0.1, seed=0):
    from manapy_amd.meshgen import jittered_triangles
    from manapy_amd.ddm import Domain
    el, nd = jittered_triangles(nx, ny, lx, ly, jitter, seed)
    return Domain(2, mesh=(el, nd))


def make_tet_domain(n, jitter=0.08, seed=0):
    from manapy_amd.meshgen import kuhn_tetrahedra
    from manapy_amd.ddm import Domain
    el, nd = kuhn_tetrahedra(n, 1.0, jitter, seed)
    return Domain(3, mesh=(el, nd))


@pytest.fixture(scope="session")
def dom2d():
    return make_tri_domain(24, 18, 1.0, 1.0, 0.1, 11)


@pytest.fixture(scope="session")
def dom3d():
    return make_tet_domain(5, 0.08, 5)
