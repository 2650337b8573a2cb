"""Shared test set-up: the `gpu` marker, the package loader and the oracle import path.

The package directory name contains hyphens, so it is loaded by path under the module name
``mibwt``.
"""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "masterarbeit_parallele_konstruktion_der_burrows-wheeler-transformation_auf_der_gpu_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def load_mibwt():
    if "mibwt" in sys.modules:
        return sys.modules["mibwt"]
    spec = importlib.util.spec_from_file_location("mibwt", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mibwt"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def mibwt():
    return load_mibwt()


@pytest.fixture(scope="session")
def storage(mibwt):
    s = mibwt.Storage(1 << 20)
    yield s
    s.close()
