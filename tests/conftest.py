"""Shared fixtures: the package (its directory name has a hyphen, so it is loaded by path), the
oracle, the golden files and the GPU marker."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_package():
    if "best_hydro_amd" in sys.modules:
        return sys.modules["best_hydro_amd"]
    d = os.path.join(ROOT, "best-hydro_amd")
    spec = importlib.util.spec_from_file_location("best_hydro_amd", os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["best_hydro_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def bh():
    return load_package()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def example_2eu7(oracle):
    return oracle.read_best(os.path.join(GOLDEN, "2eu7_02562"))


@pytest.fixture(scope="session")
def example_two(oracle):
    return oracle.read_best(os.path.join(GOLDEN, "1aki49_atom_01806"))


def no_torch_loaded():
    """Tests that capture HIP graphs (the series) are not run once torch has been imported in the
    process: with torch's own HIP libraries loaded beside the library's, an 8-mesh series capture
    ended the process with a segmentation fault (host side, no GPU fault), and in another run a
    larger 3-mesh capture did; the cause is not shown, so that state is kept out.  The one test
    that imports torch (test_zz_autograd.py) runs after them."""
    if "torch" in sys.modules:
        pytest.skip("graph capture is not run after torch's HIP libraries are loaded in this process")


def rel(a, b):
    return abs(a - b) / abs(b)


def golden_expected(name):
    """Reads one of the reference's expected_best_*.txt files copied to tests/golden as JSON."""
    import json
    with open(os.path.join(GOLDEN, "expected_best.json")) as f:
        return json.load(f)[name]
