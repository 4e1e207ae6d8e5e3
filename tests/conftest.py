"""Shared test setup: markers, import paths and the package loader."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "gpu-merkle-patricia-trie_amd")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def load_package():
    """Import gpu-merkle-patricia-trie_amd/ (not a valid identifier) as gmpt_amd."""
    if "gmpt_amd" in sys.modules:
        return sys.modules["gmpt_amd"]
    spec = importlib.util.spec_from_file_location("gmpt_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gmpt_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run on the GPU box)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gmpt():
    return load_package()


@pytest.fixture(scope="session")
def workloads(gmpt):
    return gmpt.workloads


@pytest.fixture(scope="session")
def oracle():
    import oracle as _o
    return _o
