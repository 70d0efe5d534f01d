import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PKG_NAME = "embree-benchmark-protoray_amd"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI of libprt_hip.so)")


@pytest.fixture(scope="session")
def prt():
    return importlib.import_module(PKG_NAME)


@pytest.fixture(scope="session")
def scenes():
    return importlib.import_module(PKG_NAME + ".scenes")


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def dev(prt):
    d = prt.HipDevice(0)
    yield d
    d.close()
