import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmgl.so on cuda:0)")


def load_mgl():
    """The product package lives in the hyphenated folder mg-alphagcd_amd/."""
    if "mg_alphagcd_amd" in sys.modules:
        return sys.modules["mg_alphagcd_amd"]
    pkg = os.path.join(ROOT, "mg-alphagcd_amd")
    spec = importlib.util.spec_from_file_location(
        "mg_alphagcd_amd", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    m = importlib.util.module_from_spec(spec)
    sys.modules["mg_alphagcd_amd"] = m
    spec.loader.exec_module(m)
    return m


@pytest.fixture(scope="session")
def mgl():
    return load_mgl()


@pytest.fixture(scope="session")
def orc():
    import oracle_lib
    return oracle_lib


@pytest.fixture(scope="session")
def ctx(mgl):
    c = mgl.Context(0)
    yield c
    c.close()
