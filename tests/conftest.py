"""Shared test set-up: the `gpu` marker, the package and oracle imports."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ptref  # noqa: E402  (the oracle binding; it also imports the package as ptracker_amd)

ptref.package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libptgpu kernels)")


@pytest.fixture(scope="session")
def pk():
    return ptref.package()


@pytest.fixture(scope="session")
def oracle():
    return ptref


@pytest.fixture(scope="session")
def tracker(pk):
    t = pk.Tracker(0)
    yield t
    t.close()
