import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ddc-svd_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def gpu_case(name: str) -> None:
    """Print the case about to touch the GPU, flushed, so a crash names it."""
    print(f"[gpu-case] {name}", flush=True)


@pytest.fixture(scope="session")
def ddcsvd():
    import ddcsvd as m
    m.load()
    return m


@pytest.fixture(scope="session")
def po():
    import pyoracle
    pyoracle.oracle_lib()
    return pyoracle
