import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "collocation-solver-cuda_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def load_package():
    """Import collocation-solver-cuda_amd/ under the name colloc_amd."""
    if "colloc_amd" in sys.modules:
        return sys.modules["colloc_amd"]
    spec = importlib.util.spec_from_file_location("colloc_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["colloc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcolloc.so on the device)")
