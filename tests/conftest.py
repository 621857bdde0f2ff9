import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def load_pi():
    """The package lives in a folder whose name is not an identifier; load it as `pi_amd`."""
    name = "pi_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.join(REPO, "planktonindividuals.jl_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libpi_hip.so")


@pytest.fixture(scope="session")
def pi():
    return load_pi()


@pytest.fixture(scope="session")
def orc():
    from oracle import orc as o
    o.load()
    return o
