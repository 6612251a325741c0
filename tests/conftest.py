import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.Checker("oracle")


@pytest.fixture(scope="session")
def ctx():
    import bielcc_amd
    c = bielcc_amd.Context(0)
    c.set_tuning("mixed_far32", 1)   # mixed-precision tests exercise the FP32 far sums
    yield c
    c.close()
