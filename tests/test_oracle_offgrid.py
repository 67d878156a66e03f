"""Pins of the float64 off-grid refinement oracle (oracle/offgrid_oracle.py).

Independent checks: the continuous atoms at integer positions are the
prefix-DFT columns; at noiseless data and the true positions the projected
Jacobian equals a central finite difference of the variable-projection
residual (its neglected term is proportional to the residual); the
refinement recovers planted off-grid frequencies from noiseless data; and
the whole driver matches the reference's golden runs
(tests/golden/offgrid_*.npz, written by tests/golden/make_golden.py from
torch_sbl.offgrid.refine_offgrid_fourier on CPU).  The complex128 run is
matched to rounding; the complex64 runs (the reference in float32) to 5e-5
bins in position and 5e-5 relative in amplitude.
"""

import numpy as np
import pytest
from conftest import golden_path

from oracle import offgrid_oracle as oo
from oracle.fbme_oracle import dense_prefix_matrix

F64_EPS = float(np.finfo(np.float64).eps)
F64_TINY = float(np.finfo(np.float64).tiny)


def test_atoms_at_integer_positions_are_prefix_dft_columns():
    grid, samples = (12, 10), (7, 6)
    a = dense_prefix_matrix(grid, samples)
    pos = np.array([[3.0, 7.0], [0.0, 0.0], [11.0, 9.0]])
    got = oo.continuous_dictionary(pos, grid, samples)
    idx = (pos[:, 0] * grid[1] + pos[:, 1]).astype(int)
    np.testing.assert_allclose(got, a[:, idx], atol=1e-12)


def _planted(seed=0, grid=(32, 24), samples=(20, 16), k=3):
    rng = np.random.default_rng(seed)
    pos = np.array([[5.3, 4.7], [17.8, 12.2], [26.1, 20.6]])[:k]
    amp = rng.standard_normal(k) + 1j * rng.standard_normal(k)
    y = oo.continuous_dictionary(pos, grid, samples) @ amp
    return grid, samples, pos, amp, y


def test_projected_jacobian_is_finite_difference_at_zero_residual():
    grid, samples, pos, amp, y = _planted()
    d = oo.continuous_dictionary(pos, grid, samples)
    a, chol = oo.fit_amplitudes(d, y, 0.0, F64_EPS, F64_TINY)
    jac = oo.projected_jacobian(d, a, chol, grid, samples)

    def resid(p):
        dd = oo.continuous_dictionary(p, grid, samples)
        aa, _ = oo.fit_amplitudes(dd, y, 0.0, F64_EPS, F64_TINY)
        return y - dd @ aa

    h = 1e-6
    for col in range(pos.size):
        dp = np.zeros(pos.size)
        dp[col] = h
        fd = (resid(pos + dp.reshape(pos.shape)) - resid(pos - dp.reshape(pos.shape))) / (2 * h)
        # r = y - D a  =>  dr/dp = -(projected dD/dp a)
        np.testing.assert_allclose(fd, -jac[:, col], rtol=1e-5, atol=1e-6 * np.abs(jac[:, col]).max())


def test_refinement_recovers_planted_positions():
    grid, samples, pos, amp, y = _planted(1)
    start = np.round(pos)
    out = oo.refine_single(y, start, grid, samples, max_iter=30, tolerance=1e-10, max_step=0.35,
                           damping=1e-3, ridge=0.0, eps=F64_EPS, tiny=F64_TINY)
    np.testing.assert_allclose(out["positions"], pos, atol=1e-7)
    np.testing.assert_allclose(out["amplitudes"], amp, rtol=1e-7)
    assert out["residual_norm"] < 1e-6 * np.linalg.norm(y)


@pytest.mark.parametrize("name", ["offgrid_64x64_40x40_c128", "offgrid_64x64_40x40_c64",
                                  "offgrid_64x64_40x40_b2_k5", "offgrid_128x96_80x60",
                                  "offgrid_1d_256_160_c64", "offgrid_1d_512_200_c128",
                                  "offgrid_1d_128_96_b2_k4", "offgrid_128x128_80x80_k40",
                                  "offgrid_128x128_80x80_c128_k36", "offgrid_96x96_64x64_b2_c128_k48",
                                  "offgrid_1d_1024_640_c128_k40"])
def test_offgrid_matches_reference_golden(name):
    g = np.load(golden_path(name + ".npz"))
    exact = g["y"].dtype == np.complex128
    ftype = np.float64 if exact else np.float32
    count = int(g["component_count"])
    out = oo.refine_offgrid(g["y"], g["x"], tuple(int(v) for v in g["grid"]),
                            component_count=None if count < 0 else count,
                            eps=float(np.finfo(ftype).eps), tiny=float(np.finfo(ftype).tiny))
    tol = 1e-9 if exact else 5e-5
    for i, o in enumerate(out):
        assert o["positions"].shape == g["positions"][i].shape
        np.testing.assert_allclose(o["initial_positions"], g["initial_positions"][i], atol=0)
        np.testing.assert_allclose(o["positions"], g["positions"][i], atol=tol)
        np.testing.assert_allclose(o["amplitudes"], g["amplitudes"][i], rtol=tol, atol=tol)
        np.testing.assert_allclose(o["position_std"], g["position_std"][i], rtol=1e-3 if not exact else 1e-9)
        assert o["residual_norm"] == pytest.approx(float(g["residual_norm"][i]), rel=1e-5)
        if exact:
            assert o["iterations"] == int(g["iterations"][i])
            assert o["converged"] == bool(g["converged"][i])
