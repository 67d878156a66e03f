"""Shared test setup: import paths, the `gpu` marker and GPU fixtures."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for path in (ROOT, os.path.join(ROOT, "fbme-sbl_amd")):
    if path not in sys.path:
        sys.path.insert(0, path)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X GPU and the built libfbme_hip.so")


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a GPU (run them with -m gpu on the GPU box)")
    return torch.device("cuda", 0)


def golden_path(name):
    return os.path.join(GOLDEN, name)
