"""Shared fixtures.  `-m gpu` tests need an MI355X; everything else runs on CPU."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "sgm-bench_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (run on the MI355X box)")
    config.addinivalue_line("markers", "slow: full-size oracle runs (seconds each)")


@pytest.fixture(scope="session")
def goldens():
    with open(os.path.join(ROOT, "tests", "golden", "goldens.json")) as f:
        return json.load(f)


def make_scene(spec):
    import sgm_amd
    return sgm_amd.make_scene(spec)


@pytest.fixture(scope="session")
def scene():
    return make_scene


@pytest.fixture(scope="session")
def ctx():
    """One engine context for the GPU tests (device 0)."""
    import sgm_amd
    c = sgm_amd.Context(0)
    yield c
    c.close()


def rng_image(H, W, seed):
    return np.random.default_rng(seed).integers(0, 256, size=(H, W), dtype=np.uint8)
