import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a HIP device (MI355X) and the built library")


@pytest.fixture(scope="session")
def c1_golden():
    return dict(np.load(os.path.join(GOLDEN, "c1_golden.npz")))


@pytest.fixture(scope="session")
def paths_golden():
    meta = json.load(open(os.path.join(GOLDEN, "paths_golden.json")))
    return meta, dict(np.load(os.path.join(GOLDEN, "paths_golden.npz")))


@pytest.fixture(scope="session")
def siddon_cases():
    return json.load(open(os.path.join(GOLDEN, "siddon_cases.json")))


@pytest.fixture(scope="session")
def tables_golden():
    return dict(np.load(os.path.join(GOLDEN, "tables_golden.npz")))


def have_ref():
    return os.path.exists(os.path.join(REF_DIR, "ref_driver"))


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_max(a, b):
    """max |a-b| / max |b| (relative to the image's dynamic range)."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
