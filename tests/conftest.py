"""pytest configuration: the `gpu` marker and shared helpers.

`-m "not gpu"` runs on any CPU host (oracle pins, host logic, ABI exports);
`-m gpu` runs on an MI355X and calls the HIP library through its C ABI.
"""
import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN
