import json
import os
import sys

import numpy as np
import pytest

TESTS = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(TESTS)
for p in (os.path.join(REPO, "scale-dg-kernel-extraction_amd"), os.path.join(REPO, "oracle"),
          os.path.join(TESTS, "golden"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdgk.so on the GPU)")


@pytest.fixture(scope="session")
def golden():
    import cases
    with open(cases.GOLDEN_JSON) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a GPU is visible; returns the torch module."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch


def say(msg):
    """Names the case on stdout before a GPU step runs."""
    print(msg, flush=True)


def full_field(a, NeA):
    f = np.zeros((a.shape[0], NeA), order="F")
    f[:, :a.shape[1]] = a
    return f
