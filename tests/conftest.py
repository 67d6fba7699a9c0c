import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def load_package():
    """Import the product package (directory opencfd-comb_amd/) as ``ocfd_amd``."""
    if "ocfd_amd" in sys.modules:
        return sys.modules["ocfd_amd"]
    pkg_dir = os.path.join(ROOT, "opencfd-comb_amd")
    spec = importlib.util.spec_from_file_location("ocfd_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ocfd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def ocfd():
    return load_package()


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
