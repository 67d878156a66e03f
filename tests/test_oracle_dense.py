"""Dense EM-SBL oracle (oracle/dense_em.py) against the reference's
dense_em_sbl golden vectors (complex128, 30 and 200 iterations) and the
textbook Gaussian posterior."""

import numpy as np
import pytest

from conftest import golden_path
from oracle import dense_em
from oracle import fbme_oracle as fo


@pytest.mark.parametrize("iters", [30, 200])
def test_dense_em_matches_reference_golden(iters):
    g = np.load(golden_path("general_32x32_24x24.npz"))
    out = dense_em.dense_em_sbl(g["y"].astype(np.complex128), (32, 32), iters)
    pre = f"dense{iters}_"
    xr = g[pre + "x"]
    rel = np.linalg.norm(out["x"] - xr) / np.linalg.norm(xr)
    assert rel < 1e-6
    assert abs(out["noise_variance"] / g[pre + "noise_variance"] - 1) < 1e-6
    assert abs(out["effective_dof"] - g[pre + "effective_dof"]) < 1e-6


def test_dense_posterior_is_textbook():
    rng = np.random.default_rng(0)
    a = fo.dense_prefix_matrix((6, 5), (4, 3))
    gam = rng.uniform(0.1, 2.0, a.shape[1])
    y = rng.standard_normal(a.shape[0]) + 1j * rng.standard_normal(a.shape[0])
    mean, var, _ = dense_em.dense_posterior(a, y, gam, 0.3, eps=0.0)
    cov = np.linalg.inv(np.diag(1 / gam) + np.conj(a.T) @ a / 0.3)
    np.testing.assert_allclose(mean, cov @ np.conj(a.T) @ y / 0.3, atol=1e-10)
    np.testing.assert_allclose(var, np.real(np.diag(cov)), atol=1e-10)
