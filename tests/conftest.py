import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "bvh-based-collision-detection-scheme_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_package():
    """Import the product package from its (hyphenated) directory."""
    if "bvhcd_amd" in sys.modules:
        return sys.modules["bvhcd_amd"]
    spec = importlib.util.spec_from_file_location("bvhcd_amd", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bvhcd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def bvhcd():
    return load_package()


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc  # oracle/oracle.py
    orc.lib()
    return orc
