"""The float64 oracles against reference runs with a fixed noise variance,
uniform initial variances and Hadamard probes (tests/options_golden.py):
the plan numerics (UAMP sweep, polish and Stable diagnostics) and the
general solver in mode 'cg'.  Planted slips (the noise estimated, the
backprojection blend kept) fail the comparison.  The Hadamard probes: the
oracle's (Sylvester doubling, Morton codes by loops) equal the solver's
(popcount parity), and a full basis gives Z Z^T / P = I, the exact
diagonal."""

import numpy as np
import pytest

import options_golden as og
from oracle import cg_oracle as co
from oracle import fbme_oracle as fo


def _plan(g, grid, cfg):
    out = fo.fbme_plan(g["y"][0].astype(np.complex128), grid, config=cfg)
    return out


@pytest.mark.parametrize("case", og.PLAN)
def test_plan_options_match_reference(case):
    g, grid, over = og.load(case)
    out = _plan(g, grid, og.plan_config(g, over))
    og.compare_plan(g, out)
    if "fixed_noise_variance" in over:
        fixed = np.float32(over["fixed_noise_variance"])
        assert float(out["noise_variance"]) == pytest.approx(float(fixed), rel=1e-6)


def test_plan_option_slips_fail():
    g, grid, over = og.load("options_fast_uniform_init_96x80_60x48")
    cfg = og.plan_config(g, over)
    cfg.initialization_blend = 0.1          # backprojection kept
    assert og.rel(_plan(g, grid, cfg)["x"], g["x"]) > 1e-4
    g, grid, over = og.load("options_fast_fixed_noise_96x80_60x48")
    cfg = og.plan_config(g, {})             # the noise estimated
    assert og.rel(_plan(g, grid, cfg)["x"], g["x"]) > 1e-4


# the float64 oracle against the float32 reference: the PCG stopping
# tolerance bounds the agreement.  With weak preconditioners PCG counts
# differ by one check (8 steps) at some EM iterations: rank 8 lands at
# 1.8e-4 on x after 25 EM iterations, rank 0 at 3.7e-4 on the noise.
CG_BAR = {"options_cg_rank8_48x40_40x30": (3e-4, 1e-4), "options_cg_rank0_32x32_24x24": (1e-4, 1e-3)}


@pytest.mark.parametrize("case", og.CG)
def test_cg_options_match_reference(case):
    g, grid, over = og.load(case)
    out = co.sbl_cg(g["y"][0], grid, og.cg_config(over))
    assert out["iterations"] == int(g["iterations"])
    assert out["active_set_size"] == int(g["active_set_size"])
    xbar, nbar = CG_BAR.get(case, (1e-4, 1e-4))
    assert og.rel(out["x"], g["x"]) < xbar
    assert float(out["noise_variance"]) == pytest.approx(float(g["noise_variance"]), rel=nbar)
    if "fixed_noise_variance" in over:
        fixed = float(np.float32(over["fixed_noise_variance"]))
        assert float(out["noise_variance"]) == pytest.approx(fixed, rel=1e-6)


@pytest.mark.parametrize("shape,p", [((24, 24), 32), ((40, 30), 16), ((5, 7), 11), ((4, 8), 32), ((37,), 9)])
def test_hadamard_probes_equal_oracle(shape, p):
    from torch_sbl import pcg
    import torch
    g = torch.Generator(device="cpu")
    g.manual_seed(7)
    got = pcg._hadamard_signs(g, 2, shape, p).numpy()
    ref = co.hadamard_probes(2, shape, p, 7)
    np.testing.assert_array_equal(got, ref)
    assert set(np.unique(got)) <= {-1, 1}


def test_full_rank_nystrom_is_the_exact_inverse_up_to_scale():
    """Rank M: the Nystrom approximation is A Gamma A^H itself, the missed
    trace is 0 (tr A Gamma A^H = M sum gamma), so bg = noise and the
    flattened preconditioner is (lambda_min + noise) Q^-1."""
    rng = np.random.default_rng(4)
    op = co.Operator((8, 8), (6, 4))
    gamma = rng.gamma(0.5, 1.0, (2, 64))
    noise = np.array([0.3, 0.05])
    pre = co.nystrom_preconditioner(gamma, noise, op, 24, 5, 1e-12, 1e-30)
    _, bg, u, lam = pre
    np.testing.assert_allclose(bg, noise, rtol=1e-7)     # the 1e-11 Cholesky loading moves the trace
    v = rng.standard_normal((2, 24, 3)) + 1j * rng.standard_normal((2, 24, 3))
    got = co.apply_preconditioner(v, pre)
    for b in range(2):
        q = op.q(np.eye(24, dtype=complex)[None].repeat(1, 0), gamma[b:b + 1], noise[b:b + 1])[0]
        ref = (lam[b, -1] + noise[b]) * np.linalg.solve(q, v[b])
        np.testing.assert_allclose(got[b], ref, rtol=1e-6, atol=1e-8 * np.abs(ref).max())


def test_nystrom_slips_fail():
    """The informed half taken from the weakest atoms, or the missed trace
    dropped from the background, changes the reference's PCG counts."""
    g, grid, over = og.load("options_cg_nystrom_rank12_64x64_40x40")
    good = co.sbl_cg(g["y"][0], grid, og.cg_config(over))
    assert list(map(int, good["cg_iterations"])) == g["history_cg_iterations"].astype(int).tolist()
    keep = co.nystrom_preconditioner
    try:
        def no_trace(gamma, noise, op, rank, seed, eps, tiny):
            kind, bg, u, lam = keep(gamma, noise, op, rank, seed, eps, tiny)
            return kind, np.maximum(noise, tiny), u, lam
        co.nystrom_preconditioner = no_trace
        bad = co.sbl_cg(g["y"][0], grid, og.cg_config(over))
    finally:
        co.nystrom_preconditioner = keep
    assert list(map(int, bad["cg_iterations"])) != g["history_cg_iterations"].astype(int).tolist()


def test_full_hadamard_basis_gives_the_exact_diagonal():
    """P equal to the embedding: Z = D H, so Z Z^T / P = I and the
    Hutchinson estimate of any diag(M) is exact."""
    z = co.hadamard_probes(1, (4, 8), 32, 3)[0]
    np.testing.assert_array_equal(z @ z.T / 32, np.eye(32))
    rng = np.random.default_rng(0)
    mat = rng.standard_normal((32, 32))
    est = np.mean(z * (mat @ z), axis=1)
    np.testing.assert_allclose(est, np.diag(mat), atol=1e-12)
    with pytest.raises(ValueError):
        co.hadamard_probes(1, (4, 8), 33, 3)
