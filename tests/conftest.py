import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
BASELINE = os.path.join(GOLDEN, "baseline")
POTENTIALS = os.path.join(GOLDEN, "potentials")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def dmgx():
    return importlib.import_module("dmg-md_amd")


@pytest.fixture(scope="session")
def workloads():
    return importlib.import_module("dmg-md_amd.workloads")
