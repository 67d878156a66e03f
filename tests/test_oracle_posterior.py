"""Pins of the oracle's reduced posterior and evidence (solver.py:3591-3779)
against the textbook Gaussian posterior and slogdet of the data covariance,
and of the Stable exchange's deletion gain against a brute-force evidence
difference."""

import numpy as np
import pytest

from oracle import fbme_oracle as fo


def _problem(seed, grid=(12, 10), samples=(7, 8), k=6):
    rng = np.random.default_rng(seed)
    n = grid[0] * grid[1]
    sup = np.stack([rng.choice(n, k, replace=False)])
    cols = fo.fourier_columns(sup, grid, samples)[0]
    m = cols.shape[0]
    y = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    gam = rng.uniform(0.1, 3.0, k)
    nu = 0.7
    return sup, cols, y, gam, nu, m


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_reduced_posterior_matches_textbook(seed):
    sup, cols, y, gam, nu, m = _problem(seed)
    gram = (np.conj(cols.T) @ cols)[None]
    cdat = (np.conj(cols.T) @ y)[None]
    mu, sig, dd, nle = fo.reduced_posterior(
        gram, cdat, np.array([np.vdot(y, y).real]), gam[None], np.array([nu]), m,
        eps=fo.F32_EPS, tiny=fo.F32_TINY, jitter_multiplier=0.0)
    cov = np.linalg.inv(np.diag(1 / gam) + np.conj(cols.T) @ cols / nu)
    np.testing.assert_allclose(mu[0], cov @ np.conj(cols.T) @ y / nu, rtol=1e-10)
    np.testing.assert_allclose(sig[0], np.real(np.diag(cov)), rtol=1e-10)
    c = nu * np.eye(m) + cols @ np.diag(gam) @ np.conj(cols.T)
    sign, logdet = np.linalg.slogdet(c)
    quad = np.real(np.vdot(y, np.linalg.solve(c, y)))
    assert abs(nle[0] - (logdet + quad)) < 1e-9 * abs(logdet + quad)
    np.testing.assert_allclose(dd[0], np.maximum((gam - sig[0]) / gam**2, 0),
                               rtol=1e-12)


def test_deletion_gain_is_evidence_difference():
    """-log(Sigma/gamma) - gamma |q|^2/(Sigma/gamma) equals the drop in
    -log evidence when gamma_j -> 0 (solver.py:3219-3317)."""
    sup, cols, y, gam, nu, m = _problem(5)
    gram = (np.conj(cols.T) @ cols)[None]
    cdat = (np.conj(cols.T) @ y)[None]
    e = np.array([np.vdot(y, y).real])
    kw = dict(eps=fo.F32_EPS, tiny=fo.F32_TINY, jitter_multiplier=0.0)
    mu, sig, _, nle = fo.reduced_posterior(gram, cdat, e, gam[None],
                                           np.array([nu]), m, **kw)
    gains = fo.active_deletion_gain(mu, sig, gam[None], fo.F32_TINY)[0]
    for j in range(len(gam)):
        keep = [i for i in range(len(gam)) if i != j]
        g2 = gram[:, keep][:, :, keep]
        _, _, _, nle2 = fo.reduced_posterior(g2, cdat[:, keep], e,
                                             gam[None, keep], np.array([nu]),
                                             m, **kw)
        # gain = nle(with j) - nle(without j)
        assert abs(gains[j] - (nle[0] - nle2[0])) < 1e-8 * (1 + abs(gains[j]))


def test_cholesky_loading_matches_reference_rule():
    """8 eps_f32 * mean|diag| loading changes the posterior at the 1e-6 level
    only (solver.py:2907-2920)."""
    sup, cols, y, gam, nu, m = _problem(7)
    gram = (np.conj(cols.T) @ cols)[None]
    cdat = (np.conj(cols.T) @ y)[None]
    e = np.array([np.vdot(y, y).real])
    a = fo.reduced_posterior(gram, cdat, e, gam[None], np.array([nu]), m,
                             eps=fo.F32_EPS, tiny=fo.F32_TINY, jitter_multiplier=8.0)
    b = fo.reduced_posterior(gram, cdat, e, gam[None], np.array([nu]), m,
                             eps=fo.F32_EPS, tiny=fo.F32_TINY, jitter_multiplier=0.0)
    rel = np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0])
    assert 0 < rel < 1e-4
