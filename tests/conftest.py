"""Shared test set-up: import paths, the `gpu` marker, and helpers.

The oracle (oracle/) is the checker: it is built here with gcc if needed and
only ever compared against; the product path is libeppx.so.
"""
import json
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "etherparallelparticles.jl_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libeppx kernels)")


def load_golden(name="reference_fixtures.json"):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden():
    return load_golden()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    return O


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def parallel():
    """One device context for the whole GPU session."""
    if not gpu_available():
        pytest.skip("no GPU")
    import eppx
    if not os.path.exists(eppx.library_path()):
        eppx.build()
    p = eppx.Environment.Parallel(device=0)
    p.ctx  # open now so a broken device fails here
    yield p
    p.close()
