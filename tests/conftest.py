import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle_lib():
    """The CPU oracle, built on demand (g++ only)."""
    from oracle import cpu
    if not cpu.available():
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "oracle"], check=True)
    return cpu


@pytest.fixture(scope="session")
def ref_lib():
    """The compiled reference checker; skipped where it was not built (no reference sources)."""
    from oracle import ref
    if not ref.available():
        pytest.skip("reference checker oracle/_ref not built (reference sources absent)")
    return ref


@pytest.fixture(scope="session")
def engine():
    """One engine context on cuda:0 for the GPU tests; fails loudly if the library is missing. The
    velocity correction's pressure mean is set to the reference's sequential order here, so the
    tests compare it with the reference bit for bit; the tree order has its own tests
    (test_gpu_velocity.py, tree_engine)."""
    from ding_nsch_core_solver_amd import engine as E
    eng = E.Engine(0)
    eng.set_exact_mean(True)
    yield eng
    eng.close()
