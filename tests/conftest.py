import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fortran-mpi-fdtd_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs kernels on an MI355X through the C ABI")


def relerr(a, b):
    """Normwise relative difference max|a-b| / max|b|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max()
    return float(np.abs(a - b).max() / (scale if scale > 0 else 1.0))


def check(label, err, tol):
    """Print a measured error beside its tolerance (visible with -s), then
    assert it."""
    print(f"[parity] {label}: {err:.3e} (tol {tol:.0e})", flush=True)
    assert err < tol, (label, err, tol)


@pytest.fixture(scope="session")
def golden():
    import json

    def load(name):
        z = np.load(os.path.join(GOLDEN, name + ".npz"))
        spec = json.loads(bytes(z["spec"]).decode())
        if "arg" in z:
            for d in spec["dft"]:
                d["omega"] = list(z["arg"])
        return spec, z

    return load


@pytest.fixture(scope="session")
def golden_npz():
    """A golden fixture without an embedded workload spec."""

    def load(name):
        return np.load(os.path.join(GOLDEN, name + ".npz"))

    return load
