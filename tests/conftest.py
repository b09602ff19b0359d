import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def fvm():
    return importlib.import_module("2d_fvm_cuda_amd")


@pytest.fixture(scope="session")
def problems():
    return importlib.import_module("2d_fvm_cuda_amd.problems")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu(fvm):
    """The HIP library, loaded; GPU tests fail loudly if it cannot run."""
    fvm.lib()
    assert fvm.device_available(), "no HIP device visible to libfvm2d"
    return fvm


def to_cons(U5, gamma=1.4):
    """(rho, rho u, rho v, E) planes of a [ny][nx][5] primitive array."""
    r, p, u, v = U5[..., 0], U5[..., 1], U5[..., 2], U5[..., 3]
    return np.stack([r, r * u, r * v, p / (gamma - 1.0) + 0.5 * r * (u * u + v * v)])


def rel_errors(A, B, gamma=1.4):
    """Relative L1 and L-infinity of A against reference B, per conserved
    variable (rho, rho u, rho v, E)."""
    a, b = to_cons(A, gamma), to_cons(B, gamma)
    l1 = np.abs(a - b).sum(axis=(1, 2)) / np.maximum(np.abs(b).sum(axis=(1, 2)), 1e-300)
    linf = np.abs(a - b).max(axis=(1, 2)) / np.maximum(np.abs(b).max(axis=(1, 2)), 1e-300)
    return l1, linf
