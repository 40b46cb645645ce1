"""pytest configuration: the ``gpu`` marker and import paths.

CPU runs use ``-m "not gpu"``; GPU tests call the HIP engine through its C ABI
and compare against the CPU oracle (oracle/), which only tests may import.
"""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", os.path.join("speckle-stacker_amd", "python")):
    p = os.path.join(REPO, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built engine library")


@pytest.fixture(scope="session")
def ctx():
    import smip_amd
    c = smip_amd.Context(0)
    yield c
    c.close()
