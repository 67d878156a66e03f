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
        pytest.fail("GPU tests need a GPU (run them with -m gpu on an MI355X machine)")
    return torch.device("cuda", 0)


def golden_path(name):
    return os.path.join(GOLDEN, name)


def assert_history_matches(res, g, *, upto=None, rtol=1e-4, pcg_slack=0, change_atol=2e-5):
    """Per-EM-iteration agreement of a CG solve with a reference golden run:
    identical PCG iteration counts, noise variance (batch mean) within
    ``rtol``, and relative changes ||mu_t - mu_t-1|| / ||mu_t|| within 2e-5
    absolute (a difference of two close iterates: a 1e-6 relative difference
    in mu moves it by ~1e-6 absolute, 1e-5 seen at changes of 1e-3), for
    iterations 1..upto (all when None).  ``pcg_slack`` EM iterations may
    stop their PCG one convergence check (8 iterations) apart from the
    reference: a relative residual that lands on the tolerance at a check
    is decided by float32 rounding.  Returns the largest relative noise
    deviation seen."""
    import numpy as np
    hn = np.asarray(res.history["noise_variance"], np.float64)
    rn = np.asarray(g["cg_history_noise_variance"], np.float64)
    hc = np.asarray(res.history["cg_iterations"])
    rc = np.asarray(g["cg_history_cg_iterations"])
    hr = np.asarray(res.history["relative_change"], np.float64)
    rr = np.asarray(g["cg_history_relative_change"], np.float64)
    n = min(len(hn), len(rn)) if upto is None else upto
    assert len(hn) >= n and len(rn) >= n
    off = np.flatnonzero(hc[:n] != rc[:n])
    assert off.size <= pcg_slack, ("PCG counts differ at EM iterations", (off + 1).tolist())
    assert np.all(np.abs(hc[off] - rc[off]) == 8), (hc[off], rc[off])
    dev = np.abs(hn[:n] / rn[:n] - 1.0)
    assert dev.max() < rtol, (int(dev.argmax()) + 1, float(dev.max()))
    np.testing.assert_allclose(hr[1:n], rr[1:n], rtol=1e-3, atol=change_atol)
    return float(dev.max())
