import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# The in-process bidiagonalization tests run one persistent loop kernel per simulated peer, each
# on its own stream, and the loops wait on each other; two streams sharing one hardware queue
# would serialize them. Give the test process more hardware queues than it has such streams
# (set before the HIP runtime starts; the runtime's default is 4).
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU oracle cases")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def hip():
    """The HIP binding; fails (does not skip) when the library or the GPU is missing."""
    from excalibur_amd import hip as H
    assert H.device_count() >= 1, "no GPU visible to the HIP runtime"
    return H


def random_bidiagonal(rng, n, kind="gauss"):
    if kind == "gauss":
        return rng.standard_normal(n), rng.standard_normal(n - 1)
    if kind == "positive":
        return rng.uniform(0.5, 2.0, n), rng.uniform(0.1, 1.0, n - 1)
    if kind == "graded":
        a = 10.0 ** (-rng.uniform(0, 8, n)); b = 10.0 ** (-rng.uniform(0, 8, n - 1))
        return a, b
    if kind == "clustered":
        a = 1.0 + 1e-9 * rng.standard_normal(n); b = 1e-9 * rng.standard_normal(n - 1)
        return a, b
    raise ValueError(kind)
