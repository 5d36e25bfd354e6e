"""Shared fixtures.  GPU tests carry @pytest.mark.gpu; everything else runs on
the CPU (oracle known answers, host logic, the C-ABI library's exports)."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long oracle runs, opt in with PM_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("PM_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow oracle run (set PM_SLOW=1)")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def known():
    with open(os.path.join(GOLDEN, "reference_runs.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ref_fields():
    """Fields, codes and rank maps written by the reference binary
    (tests/golden/make_reference_fields.py)."""
    with np.load(os.path.join(GOLDEN, "reference_fields.npz")) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="session")
def ref_meta():
    with open(os.path.join(GOLDEN, "reference_fields.json")) as f:
        return json.load(f)


def ref_run(meta, group, tag):
    for r in meta[group]["runs"]:
        if r["tag"] == tag:
            return r
    raise KeyError(tag)


@pytest.fixture(scope="session")
def channel_raw():
    return np.fromfile(os.path.join(GOLDEN, "rect_channel_24_14_50_vs_5e-05.raw"),
                       dtype=np.uint8).reshape(50, 14, 24)


def random_geometry(shape, phi, seed):
    """Small random solid/fluid box (independent voxels) for parity checks."""
    rng = np.random.default_rng(seed)
    return (rng.random(shape) > phi).astype(np.uint8)
