"""Shared test setup.  Float64 results are selected before the package is imported (as the
reference suite does, ``tests/conftest.py:1-5``) so finite differences and tight parity bounds
are meaningful."""

import importlib
import os
import sys
from pathlib import Path

os.environ.setdefault("ADTLERT_ENABLE_FLOAT64", "1")

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402
import pytest  # noqa: E402
import torch  # noqa: E402

adt = importlib.import_module("ad-tlert_amd")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (HIP) and the built gfx950 library")


def pytest_collection_modifyitems(config, items):
    if torch.cuda.is_available():
        return
    skip = pytest.mark.skip(reason="requires a HIP GPU")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


def relmax(actual, expected) -> float:
    """Max-norm relative error (the reference suite's ``assert_close`` measure)."""

    actual = np.asarray(actual, dtype=float)
    expected = np.asarray(expected, dtype=float)
    return float(np.abs(actual - expected).max() / max(np.abs(expected).max(), 1e-300))
