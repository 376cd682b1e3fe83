import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hf-rs_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (os.path.join(ROOT, "oracle"), PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long CPU test")


def load_package():
    """Import hf-rs_amd/ as the package `hf_rs_amd` (its directory name has a hyphen)."""
    if "hf_rs_amd" in sys.modules:
        return sys.modules["hf_rs_amd"]
    spec = importlib.util.spec_from_file_location("hf_rs_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hf_rs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
