"""Shared test setup: the `gpu` marker, in-tree builds when missing, loaders."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden")
REFDATA = os.path.join(GOLDEN, "reference")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def _ensure_built():
    lib = os.path.join(ROOT, "nraps_amd", "libnraps.so")
    if not os.path.exists(lib):
        from nraps_amd import build
        build.build()
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(orc):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)


_ensure_built()


def has_gpu():
    import nraps_amd
    st, _ = nraps_amd.device_info()
    return st == 0


@pytest.fixture(scope="session")
def gpu():
    if not has_gpu():
        pytest.fail("no HIP device: GPU tests must run on the MI355X box")
    return True


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib
