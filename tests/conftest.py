import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_package():
    """Import the ``jax-fem_amd`` directory as the package ``jax_fem_amd``."""
    if "jax_fem_amd" in sys.modules:
        return sys.modules["jax_fem_amd"]
    pkg_dir = os.path.join(ROOT, "jax-fem_amd")
    spec = importlib.util.spec_from_file_location(
        "jax_fem_amd", os.path.join(pkg_dir, "__init__.py"), submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["jax_fem_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def jfa():
    return load_package()
