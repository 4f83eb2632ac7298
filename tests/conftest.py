"""Shared test setup: markers and paths.

Tests marked ``gpu`` need an MI355X and the built HIP library; everything
else runs on the CPU (oracle restatement, golden fixtures, host logic and the
C-ABI export check)."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "oracle", "aurora-pic_amd"):
    p = os.path.join(REPO, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import oracle_py
    return oracle_py.Oracle()


@pytest.fixture(scope="session")
def reference():
    import oracle_py
    if not oracle_py.reference_available():
        pytest.skip("reference library oracle/_ref not built (reference tree absent)")
    return oracle_py.Reference()


GOLDEN = os.path.join(REPO, "tests", "golden")


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
