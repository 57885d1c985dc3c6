"""Shared fixtures: the engine package (directory mld2p4-2_amd/) and the
oracle (test infrastructure).  GPU tests are marked `gpu`."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def load_engine():
    name = "mld2p4_amd"
    if name in sys.modules:
        return sys.modules[name]
    path = os.path.join(ROOT, "mld2p4-2_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(name, path, submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    load_engine()  # importable as mld2p4_amd (the library itself loads lazily)


@pytest.fixture(scope="session")
def eng():
    return load_engine()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle
