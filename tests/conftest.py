"""Shared test setup: markers, library handles, fixture loading.

CPU tests (-m "not gpu") check the oracle against the reference's golden
vectors and closed forms, the host-side API and the exported symbol set.
GPU tests (-m gpu) compare the HIP library with the oracle and the fixtures
through the C ABI."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "cfd-bench_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

import nusif_api as A  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


def _have_gpu():
    try:
        lib = A.load(A.gpu_lib_path("rb"))
        return lib.nusifDeviceName(None, 0) > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu_ok():
    if not os.path.exists(A.gpu_lib_path("rb")):
        pytest.fail("GPU library not built: run python cfd-bench_amd/build.py")
    if not _have_gpu():
        pytest.fail("no GPU visible to the HIP library")
    return True


_LIBS = {}


def lib(kind):
    """Cached ctypes handle: 'gpu_rb', 'gpu_mg', 'orc_rb', 'orc_mg', 'orc_gmg',
    'ref_rb', 'ref_mg'."""
    if kind not in _LIBS:
        src, solver = kind.split("_", 1)
        path = {"gpu": A.gpu_lib_path, "orc": A.oracle_lib_path, "ref": A.ref_lib_path}[src](solver)
        if not os.path.exists(path):
            if src == "ref":
                pytest.skip("reference build absent (oracle/_ref needs the reference tree)")
            pytest.fail("%s not built (python -c 'import __graft_entry__ as g; g.build()')" % path)
        _LIBS[kind] = A.load(path)
    return _LIBS[kind]


@pytest.fixture(scope="session")
def kat():
    return np.load(os.path.join(ROOT, "tests", "golden", "kat.npz"))


@pytest.fixture(scope="session")
def traj():
    return np.load(os.path.join(ROOT, "tests", "golden", "traj.npz"))
