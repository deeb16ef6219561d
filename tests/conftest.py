import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


def load_package():
    """Import the product package from cahn-hilliard_amd/ as `cahn_hilliard_amd`."""
    if "cahn_hilliard_amd" in sys.modules:
        return sys.modules["cahn_hilliard_amd"]
    pkg_dir = os.path.join(REPO, "cahn-hilliard_amd")
    spec = importlib.util.spec_from_file_location(
        "cahn_hilliard_amd", os.path.join(pkg_dir, "__init__.py"), submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cahn_hilliard_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def ch():
    return load_package()
