"""Test configuration: import paths, the `gpu` marker, shared fixtures.

`-m "not gpu"` runs everywhere (oracle checks, host logic, ABI); `-m gpu`
needs an MI355X and exercises the HIP kernels through the C-ABI.  GPU tests
never skip when the device or the library is missing: they fail.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "perf-libs-sparse_amd"))
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    with open(ROOT / "tests" / "golden" / "reference_cases.json") as f:
        return json.load(f)


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def pls():
    import plsparse

    plsparse.lib()
    return plsparse


@pytest.fixture(scope="session")
def gen():
    from plsparse import workloads

    workloads.lib()
    return workloads


def rel_err(got, ref) -> float:
    """Normwise relative error ||got - ref||_2 / ||ref||_2 (0/0 -> 0)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.linalg.norm(ref)
    num = np.linalg.norm(got - ref)
    if den == 0.0:
        return float(num)
    return float(num / den)
