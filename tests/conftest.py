"""Shared fixtures: marker registration and loaders for the engine and the oracle."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU oracle run (seconds to a minute)")


def load_engine():
    """Import the engine package from its hyphenated directory."""
    if "litter_physics_amd" in sys.modules:
        return sys.modules["litter_physics_amd"]
    pkg = os.path.join(ROOT, "litter-physics_amd")
    spec = importlib.util.spec_from_file_location("litter_physics_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["litter_physics_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def engine():
    return load_engine()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O
