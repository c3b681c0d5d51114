"""Test configuration: the `gpu` marker and shared fixtures.

CPU tests (-m "not gpu") pin the oracle against the golden vectors and the
reference build, check closed forms, the host-side API and that the C-ABI
library exports every symbol declared in include/rcp_amd.h.  GPU tests
(-m gpu) call the HIP engine through the C ABI and compare with the oracle.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP engine)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    return O


@pytest.fixture(scope="session")
def ref(oracle):
    if not oracle.have_ref():
        pytest.skip("reference build (oracle/_ref) not available")
    return oracle


def load_build_recipe():
    """The package's build recipe, loaded by path: importing the package
    fails until its native module exists."""
    import importlib.util
    path = os.path.join(ROOT, "rcpgenerator_amd", "_build.py")
    spec = importlib.util.spec_from_file_location("rcpgenerator_amd_build", path)
    recipe = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(recipe)
    return recipe


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden():
    return load_golden
