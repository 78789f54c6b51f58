import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libmlhip.so")


def load_pkg():
    return importlib.import_module("mushy-layer_amd")


def load_workloads():
    return importlib.import_module("mushy-layer_amd.workloads")


@pytest.fixture(scope="session")
def oracle():
    from oracle import mlref
    mlref.lib()
    return mlref


@pytest.fixture(scope="session")
def ml():
    return load_pkg()


@pytest.fixture(scope="session")
def wl():
    return load_workloads()


@pytest.fixture(scope="session")
def gpu_ctx(ml):
    ctx = ml.Context(0)
    yield ctx
    ctx.close()
