"""Pins of the oracle's prefix DFT operator against closed forms.

A = fft2(.)[:M1, :M2] and A^H = N ifft2(pad(.)) are checked against the
brute-force DFT submatrix, the adjoint identity, A A^H = N I and the
Dirichlet-kernel Gram closed form (torch_sbl/operators.py:237-280, 444-466).
"""

import numpy as np
import pytest

from oracle import fbme_oracle as fo

SHAPES = [((8, 8), (5, 6)), ((12, 10), (7, 10)), ((6, 15), (6, 4)),
          ((1, 16), (1, 9))]


@pytest.mark.parametrize("grid,samples", SHAPES)
def test_forward_adjoint_match_dense_matrix(grid, samples):
    rng = np.random.default_rng(0)
    a = fo.dense_prefix_matrix(grid, samples)
    x = rng.standard_normal((3, a.shape[1])) + 1j * rng.standard_normal((3, a.shape[1]))
    v = rng.standard_normal((3, a.shape[0])) + 1j * rng.standard_normal((3, a.shape[0]))
    np.testing.assert_allclose(fo.prefix_forward(x, grid, samples), x @ a.T,
                               rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(fo.prefix_adjoint(v, grid, samples),
                               v @ np.conj(a), rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("grid,samples", SHAPES)
def test_rows_orthogonal_and_adjoint_identity(grid, samples):
    a = fo.dense_prefix_matrix(grid, samples)
    n = grid[0] * grid[1]
    np.testing.assert_allclose(a @ np.conj(a.T), n * np.eye(a.shape[0]),
                               atol=1e-9)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, n)) + 1j * rng.standard_normal((1, n))
    v = rng.standard_normal((1, a.shape[0])) + 1j * rng.standard_normal((1, a.shape[0]))
    lhs = np.vdot(v, fo.prefix_forward(x, grid, samples))
    rhs = np.vdot(fo.prefix_adjoint(v, grid, samples), x)
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


@pytest.mark.parametrize("grid,samples", SHAPES)
def test_columns_and_gram_closed_form(grid, samples):
    a = fo.dense_prefix_matrix(grid, samples)
    rng = np.random.default_rng(2)
    n = grid[0] * grid[1]
    idx = np.stack([rng.choice(n, size=min(5, n), replace=False) for _ in range(2)])
    cols = fo.fourier_columns(idx, grid, samples)
    np.testing.assert_allclose(cols, np.stack([a[:, i] for i in idx]), atol=1e-10)
    gram = np.einsum("bmk,bml->bkl", np.conj(cols), cols)
    np.testing.assert_allclose(fo.gram_closed_form(idx, grid, samples), gram,
                               atol=1e-9)


def test_gram_apply_is_adjoint_of_sparse_forward():
    grid, samples = (10, 12), (6, 7)
    rng = np.random.default_rng(3)
    sup = np.stack([rng.choice(120, 6, replace=False)])
    mu = rng.standard_normal((1, 6)) + 1j * rng.standard_normal((1, 6))
    x = np.zeros((1, 120), complex)
    x[0, sup[0]] = mu[0]
    ref = fo.prefix_adjoint(fo.prefix_forward(x, grid, samples), grid, samples)
    np.testing.assert_allclose(fo._gram_apply(mu, sup, grid, samples), ref,
                               atol=1e-9)


def test_float32_phase_mimic_has_reference_sized_error():
    """The reference's float32 phase (no mod-N reduction) errs by ~1e-4 rad
    at 256^2 / 160^2; the exact columns do not."""
    grid, samples = (256, 256), (160, 160)
    idx = np.array([[65535, 40000, 255 * 256 + 3]])
    exact = fo.fourier_columns(idx, grid, samples)
    mimic = fo.fourier_columns(idx, grid, samples, phase_dtype=np.float32)
    err = np.max(np.abs(exact - mimic))
    assert 1e-6 < err < 1e-3
