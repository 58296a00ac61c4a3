import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402

sf = entry.load_package()
from oracle import sf_oracle  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libstable-fluids.so)")


_HAS_GPU = None


def has_gpu():
    global _HAS_GPU
    if _HAS_GPU is None:
        try:
            _HAS_GPU = sf.hip.device_count() > 0
        except Exception:
            _HAS_GPU = False
    return _HAS_GPU


def pytest_collection_modifyitems(config, items):
    if has_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def pkg():
    return sf


@pytest.fixture(scope="session")
def abi():
    return sf.abi


@pytest.fixture(scope="session")
def oracle():
    sf_oracle.load(sf.abi)
    return sf_oracle
