import importlib
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbcs HIP kernels)")


@pytest.fixture(scope="session")
def bcs():
    return importlib.import_module("simulation-server_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle_py

    return oracle_py


def ref_dir():
    """Outputs of the reference's own table probe (oracle/_ref), if built."""
    d = ROOT / "oracle" / "_ref"
    return d if (d / "ref_tables.json").exists() else None


def soa(a, n):
    return np.asarray(a).reshape(3, n)


def assert_bits_equal(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if a.dtype.kind == "f":
        same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
    else:
        same = a == b
    if not same.all():
        bad = np.flatnonzero(~same)
        i = bad[0]
        raise AssertionError(f"{what}: {bad.size} of {a.size} differ; first at {i}: {a.flat[i]!r} vs {b.flat[i]!r}")
