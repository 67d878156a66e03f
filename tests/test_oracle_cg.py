"""Pins of the float64 CG oracle (oracle/cg_oracle.py).

Each piece is checked against an independent computation: the Toeplitz
operator against the dense Q = noise I + A Gamma A^H, PCG against a dense
solve, the low-rank preconditioner against the dense Woodbury inverse, the
Hutchinson estimate against the exact diagonal (unbiasedness over many
probes), and the whole solver against the reference's own golden run
(tests/golden/general_32x32_24x24.npz, written by
tests/golden/make_golden.py from torch_sbl.sbl_2d(mode='cg') on CPU).
"""

from __future__ import annotations

import numpy as np
import pytest
from conftest import golden_path

from oracle import cg_oracle as co
from oracle.fbme_oracle import F32_EPS, F32_TINY, dense_prefix_matrix, fourier_columns


def _problem(seed=0, grid=(12, 10), samples=(7, 6), batch=2):
    rng = np.random.default_rng(seed)
    n, m = grid[0] * grid[1], samples[0] * samples[1]
    gamma = rng.gamma(0.5, 1.0, (batch, n))
    noise = rng.uniform(0.05, 0.2, batch)
    a = dense_prefix_matrix(grid, samples)
    return rng, gamma, noise, a, co.Operator(grid, samples)


def test_operator_matches_dense_covariance():
    rng, gamma, noise, a, op = _problem()
    d = rng.standard_normal((2, op.m, 3)) + 1j * rng.standard_normal((2, op.m, 3))
    got = op.q(d, gamma, noise)
    for b in range(2):
        q = noise[b] * np.eye(op.m) + (a * gamma[b]) @ a.conj().T
        np.testing.assert_allclose(got[b], q @ d[b], rtol=1e-12, atol=1e-12)


def test_preconditioner_is_dense_woodbury_inverse():
    rng, gamma, noise, a, op = _problem(1)
    bg, factors, inv = co.low_rank_preconditioner(gamma, noise, op, 8, 0.0, F32_TINY)
    v = rng.standard_normal((2, op.m, 2)) + 1j * rng.standard_normal((2, op.m, 2))
    got = co.apply_preconditioner(v, (bg, factors, inv))
    for b in range(2):
        order = np.argsort(-gamma[b], kind="stable")[:8]
        cols = a[:, order]
        approx = bg[b] * np.eye(op.m) + (cols * gamma[b, order]) @ cols.conj().T
        np.testing.assert_allclose(got[b], np.linalg.solve(approx, v[b]), rtol=1e-9, atol=1e-11)
        omitted = gamma[b].sum() - gamma[b, order].sum()
        assert bg[b] == pytest.approx(noise[b] + omitted, rel=1e-12)


def test_fourier_factor_columns_are_dense_columns():
    _, gamma, _, a, op = _problem(2)
    idx = np.array([[0, 5, 17, 119]])
    np.testing.assert_allclose(fourier_columns(idx, op.grid, op.samples)[0], a[:, idx[0]], atol=1e-12)


@pytest.mark.parametrize("fixed", [0, 12])
def test_pcg_solves_dense_system(fixed):
    rng, gamma, noise, a, op = _problem(3)
    rhs = rng.standard_normal((2, op.m, 4)) + 1j * rng.standard_normal((2, op.m, 4))
    x, it = co.pcg(rhs, gamma, noise, op, x0=None, tol=1e-10, max_iter=200, check=8, rank=8,
                   fixed=fixed, eps=F32_EPS, tiny=F32_TINY)
    for b in range(2):
        q = noise[b] * np.eye(op.m) + (a * gamma[b]) @ a.conj().T
        exact = np.linalg.solve(q, rhs[b])
        err = np.linalg.norm(x[b] - exact) / np.linalg.norm(exact)
        assert err < (1e-8 if fixed == 0 else 1e-2)
    if fixed:
        assert it == fixed


def test_pcg_warm_start_from_solution_stops_at_first_check():
    rng, gamma, noise, a, op = _problem(4)
    rhs = rng.standard_normal((2, op.m, 2)) + 1j * rng.standard_normal((2, op.m, 2))
    x, _ = co.pcg(rhs, gamma, noise, op, x0=None, tol=1e-12, max_iter=300, check=8, rank=8)
    _, it = co.pcg(rhs, gamma, noise, op, x0=x, tol=1e-6, max_iter=300, check=8, rank=8)
    assert it == 8


def test_hutchinson_is_unbiased_for_posterior_diagonal():
    """mean_p Re(A^H Q^-1 p . conj(A^H p)) -> diag(A^H Q^-1 A) as P grows."""
    rng, gamma, noise, a, op = _problem(5, batch=1)
    q = noise[0] * np.eye(op.m) + (a * gamma[0]) @ a.conj().T
    exact = np.real(np.sum(a.conj() * np.linalg.solve(q, a), 0))
    count = 4000
    probes = rng.choice([-1.0, 1.0], (1, op.m, count)).astype(np.complex128)
    sol = np.linalg.solve(q, probes[0])[None]
    est = np.mean(np.real(op.adj(sol) * np.conj(op.adj(probes))), 2)[0]
    # Rademacher estimator variance ~ sum_j |off-diagonal|^2 / P
    assert np.max(np.abs(est - exact) / exact) < 0.15
    assert abs(est.mean() / exact.mean() - 1) < 0.02


def test_probe_count_rule():
    cfg = co.CGConfig()
    assert co.probe_count(1024, 576, cfg) == 32               # 8 sqrt(1.78) = 10.7 -> 32
    assert co.probe_count(65536, 1024, cfg) == 128             # 2 * 64 = 128
    assert co.probe_count(65536, 25600, cfg) == 32
    assert co.probe_count(65536, 2048, cfg) == 64              # max(8 sqrt 32, 2 * 32)
    cfg.auto_probe_scaling = False
    assert co.probe_count(65536, 1024, cfg) == 32


def test_rademacher_probes_are_signs():
    p = co.rademacher_probes(2, 50, 8, 0)
    assert p.shape == (2, 50, 8)
    assert set(np.unique(p)) == {-1.0, 1.0}
    np.testing.assert_array_equal(p, co.rademacher_probes(2, 50, 8, 0))


def test_schedule():
    cfg = co.CGConfig()
    assert co.cg_schedule(1, cfg) == 1e-2
    assert co.cg_schedule(6, cfg) == pytest.approx(np.sqrt(1e-7))
    assert co.cg_schedule(16, cfg) == 1e-5


def test_sbl_cg_matches_reference_golden_run():
    g = np.load(golden_path("general_32x32_24x24.npz"))
    out = co.sbl_cg(g["y"], (32, 32))
    assert out["iterations"] == int(g["cg_iterations"])
    assert out["active_set_size"] == int(g["cg_active_set_size"])
    ref = g["cg_x"].reshape(-1)
    rel = np.linalg.norm(out["x"].reshape(-1) - ref) / np.linalg.norm(ref)
    assert rel < 1e-4
    assert out["noise_variance"] == pytest.approx(float(g["cg_noise_variance"]), rel=1e-3)
    assert out["effective_dof"] == pytest.approx(float(g["cg_effective_dof"]), rel=1e-3)
    np.testing.assert_allclose(out["cg_iterations"][:29], g["cg_history_cg_iterations"][:29], atol=8)


def test_masked_operator_is_dense_row_restriction():
    """With a mask, Q acts on the observed rows of A only."""
    rng, gamma, noise, a, _ = _problem(6)
    mask = rng.random((7, 6)) < 0.6
    op = co.Operator((12, 10), (7, 6), mask)
    ar = a[mask.reshape(-1)]
    assert op.m == int(mask.sum())
    d = rng.standard_normal((2, op.m, 2)) + 1j * rng.standard_normal((2, op.m, 2))
    got = op.q(d, gamma, noise)
    for b in range(2):
        q = noise[b] * np.eye(op.m) + (ar * gamma[b]) @ ar.conj().T
        np.testing.assert_allclose(got[b], q @ d[b], rtol=1e-12, atol=1e-12)
    bg, factors, inv = co.low_rank_preconditioner(gamma, noise, op, 8, 0.0, F32_TINY)
    order = np.argsort(-gamma[0], kind="stable")[:8]
    np.testing.assert_allclose(factors[0], ar[:, order] * np.sqrt(gamma[0, order]), atol=1e-12)


def test_full_mask_is_unmasked_solver():
    g = np.load(golden_path("general_32x32_24x24.npz"))
    cfg = co.CGConfig(max_iter=6)
    a = co.sbl_cg(g["y"], (32, 32), cfg)
    b = co.sbl_cg(g["y"], (32, 32), cfg, mask=np.ones((24, 24), bool))
    np.testing.assert_allclose(b["x"], a["x"], rtol=1e-12, atol=1e-14)
    assert b["noise_variance"] == pytest.approx(a["noise_variance"], rel=1e-12)


def test_mask_validation():
    with pytest.raises(ValueError):
        co.Operator((8, 8), (4, 4), np.ones((4, 5), bool))
    with pytest.raises(ValueError):
        co.Operator((8, 8), (4, 4), np.zeros((4, 4), bool))


@pytest.mark.parametrize("name,grid,kw", [
    ("masked_32x32_24x24", (32, 32), {}),
    ("masked_64x64_40x40_b2", (64, 64), dict(hutchinson_probes=16, auto_probe_scaling=False,
                                             cg_tolerance=1e-4, max_iter=30)),
])
def test_masked_sbl_cg_matches_reference_golden_run(name, grid, kw):
    """sbl_2d(mode='cg', observation_mask=mask) of the reference
    (solver.py:986-1024, operators.py:57-75) on 67% / 58% observed samples."""
    g = np.load(golden_path(name + ".npz"))
    out = co.sbl_cg(g["y"], grid, co.CGConfig(**kw), mask=g["mask"])
    assert out["iterations"] == int(g["cg_iterations"])
    assert out["active_set_size"] == int(g["cg_active_set_size"])
    ref = g["cg_x"].reshape(out["x"].shape)
    assert np.linalg.norm(out["x"] - ref) / np.linalg.norm(ref) < 1e-4
    np.testing.assert_allclose(out["noise_variance"], g["cg_noise_variance"], rtol=1e-3)
    np.testing.assert_allclose(out["effective_dof"], g["cg_effective_dof"], rtol=1e-3)
    np.testing.assert_allclose(out["cg_iterations"], g["cg_history_cg_iterations"], atol=8)


def test_shared_posterior_is_dense_multi_rhs_solve():
    """The joint-noise posterior (one system, snapshots and probes as
    right-hand sides) against the dense solve with the same probes."""
    rng, gamma, noise, a, op = _problem(9, batch=3)
    gamma = np.repeat(gamma[:1], 3, 0)
    noise = np.full(3, noise[0])
    y = rng.standard_normal((3, op.m)) + 1j * rng.standard_normal((3, op.m))
    probes = co.rademacher_probes(1, op.m, 64, 0).astype(np.complex128)
    mean, var, dd, _, _ = co._posterior_shared(y, gamma, noise, probes, op.adj(probes), op, None,
                                               co.CGConfig(cg_warm_start=False), 1e-12, F32_EPS, F32_TINY)
    q = noise[0] * np.eye(op.m) + (a * gamma[0]) @ a.conj().T
    np.testing.assert_allclose(mean, gamma[:1] * (a.conj().T @ np.linalg.solve(q, y.T)).T, rtol=1e-8, atol=1e-10)
    lifted = a.conj().T @ np.linalg.solve(q, probes[0])
    want = np.maximum(np.mean(np.real(lifted * np.conj(a.conj().T @ probes[0])), 1), 0.0)
    np.testing.assert_allclose(dd[0], want, rtol=1e-8, atol=1e-12)
    assert np.array_equal(dd[0], dd[2]) and np.array_equal(var[0], var[1])


def test_joint_requires_two_snapshots():
    g = np.load(golden_path("general_32x32_24x24.npz"))
    with pytest.raises(ValueError):
        co.sbl_cg(g["y"], (32, 32), co.CGConfig(max_iter=2), joint=True)


@pytest.mark.parametrize("name,grid,kw", [
    ("joint_32x32_24x24_b3", (32, 32), {}),
    ("joint_32x32_24x24_b3_noise", (32, 32), {}),
    ("joint_64x64_40x40_b4", (64, 64), dict(hutchinson_probes=16, auto_probe_scaling=False,
                                            cg_tolerance=1e-4, max_iter=30)),
])
def test_joint_sbl_cg_matches_reference_golden_run(name, grid, kw):
    """sbl_2d(mode='cg', joint_sparsity=True) of the reference, with shared
    noise (one multi-right-hand-side system) and per-snapshot noise."""
    g = np.load(golden_path(name + ".npz"))
    out = co.sbl_cg(g["y"], grid, co.CGConfig(**kw), joint=True,
                    joint_noise=bool(g["joint_noise_variance"]))
    assert out["iterations"] == int(g["cg_iterations"])
    ref = g["cg_x"].reshape(out["x"].shape)
    assert np.linalg.norm(out["x"] - ref) / np.linalg.norm(ref) < 1e-3
    np.testing.assert_allclose(out["noise_variance"], g["cg_noise_variance"], rtol=1e-3)
    np.testing.assert_allclose(out["effective_dof"], g["cg_effective_dof"], rtol=1e-3)
    np.testing.assert_allclose(out["cg_iterations"], g["cg_history_cg_iterations"], atol=8)


ACTIVE = ["active_64x64_40x40_k60", "active_48x48_32x32_k50", "active_64x64_40x40_k45",
          "active_64x64_40x40_k60_max16", "active_64x64_40x40_k45_min48"]


@pytest.mark.parametrize("case", ACTIVE)
def test_oracle_active_set_sizes_match_reference(case):
    """The reference's active-set switch (solver.py:3184-3216) at 58, 161 and
    37 atoms, with active_set_max_size=16 (never switches) and
    active_set_min_size=48: the oracle takes the same switch at the same
    iteration and matches x."""
    g = np.load(golden_path(case + ".npz"))
    over = {k[len("override_"):]: g[k].item() for k in g.files if k.startswith("override_")}
    cfg = co.CGConfig(**over)
    grid = tuple(int(v) for v in g["grid"])
    trace = []
    out = co.sbl_cg(g["y"], grid, cfg, trace=trace)
    hist = [t["active"] for t in trace]
    np.testing.assert_array_equal(hist, g["cg_history_active_set_size"][:len(hist)])
    assert out["active_set_size"] == int(g["cg_active_set_size"])
    assert out["iterations"] == int(g["cg_iterations"])
    x, xr = out["x"].reshape(-1), g["cg_x"].reshape(-1)
    rel = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    print(case, "active", out["active_set_size"], "rel", rel)
    assert rel < 1e-4    # achieved 2.8e-6 .. 1.1e-5


PRECOND = ["precond_circulant_32x32_24x24", "precond_circulant_64x64_40x40",
           "precond_circulant_masked_64x64_40x40", "precond_none_32x32_24x24"]


@pytest.mark.parametrize("case", PRECOND)
def test_oracle_preconditioners_match_reference(case):
    """cg_preconditioner='circulant' (the Strang BCCB approximation,
    operators.py:373-416) and 'none' against the reference's runs."""
    g = np.load(golden_path(case + ".npz"))
    over = {k[len("override_"):]: g[k].item() for k in g.files if k.startswith("override_")}
    cfg = co.CGConfig(**over)
    grid = tuple(int(v) for v in g["grid"])
    out = co.sbl_cg(g["y"], grid, cfg, mask=g["mask"] if "mask" in g.files else None)
    x, xr = out["x"].reshape(-1), g["cg_x"].reshape(-1)
    rel = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    print(case, "pcg", out["cg_iterations"][:8], list(g["cg_history_cg_iterations"][:8]), "rel", rel)
    assert out["iterations"] == int(g["cg_iterations"])
    assert rel < 1e-3


def test_circulant_preconditioner_is_the_strang_approximation():
    """The eigenvalues are those of the circulant matrix whose first column
    holds Q's lags (k <= M/2 ? k : k - M): built densely on a small grid and
    diagonalised, against the FFT form; the apply inverts it."""
    grid, samples = (12, 10), (7, 6)
    rng = np.random.default_rng(9)
    gamma = rng.exponential(size=(1, 120))
    noise = np.array([0.3])
    op = co.Operator(grid, samples)
    _, eig, _ = co.circulant_preconditioner(gamma, noise, op, F32_EPS)
    a = dense_prefix_matrix(grid, samples)
    q = (a * gamma[0]) @ a.conj().T + noise[0] * np.eye(42)
    m1, m2 = samples
    # Strang: wrap Q's first-column lags onto the 7 x 6 torus
    col = np.zeros((m1, m2), complex)
    for k1 in range(m1):
        for k2 in range(m2):
            d1 = k1 if k1 <= m1 // 2 else k1 - m1
            d2 = k2 if k2 <= m2 // 2 else k2 - m2
            if d1 >= 0 and d2 >= 0:
                col[k1, k2] = q[d1 * m2 + d2, 0]
            else:
                # lag d of Q = A Gamma A^H: q[(m + d), m] for a row m where both are inside
                r1, r2 = max(-d1, 0), max(-d2, 0)
                col[k1, k2] = q[(r1 + d1) * m2 + (r2 + d2), r1 * m2 + r2]
    want = np.real(np.fft.fft2(col))
    want = np.maximum(want, max(F32_EPS * np.abs(want).max(), noise[0]))
    np.testing.assert_allclose(eig[0], want, rtol=1e-10, atol=1e-10 * np.abs(want).max())
    v = rng.standard_normal((1, 42, 2)) + 1j * rng.standard_normal((1, 42, 2))
    z = co.apply_preconditioner(v, ("circulant", eig, op))
    back = np.fft.ifft2(np.fft.fft2(z.reshape(1, m1, m2, 2), axes=(1, 2)) * eig[..., None], axes=(1, 2))
    np.testing.assert_allclose(back.reshape(1, 42, 2), v, atol=1e-10)


# ---------------------------------------------------------------------------
# grid-diagonal background of the learned solver's preconditioner
# ---------------------------------------------------------------------------

def test_grid_diagonal_background_is_exact_inverse_on_the_full_grid():
    """M = N: A = F with F F^H = N I, so A diag(1 / (N (noise + N g))) A^H
    is (noise I + A diag(g) A^H)^-1 for any variances g.  A dropped factor
    N, or M in place of N, fails this."""
    grid = samples = (6, 5)
    rng = np.random.default_rng(11)
    n = 30
    gamma = rng.gamma(0.5, 1.0, (1, n))
    noise = np.array([0.07])
    a = dense_prefix_matrix(grid, samples)
    none = np.zeros((1, 0), dtype=np.int64)
    w = co.grid_diagonal_weights(gamma, noise, none)
    binv = (a * w[0]) @ a.conj().T
    q = noise[0] * np.eye(n) + (a * gamma[0]) @ a.conj().T
    np.testing.assert_allclose(binv @ q, np.eye(n), atol=1e-10)


def test_grid_diagonal_preconditioner_is_dense_inverse_of_b_plus_ffh():
    rng, gamma, noise, a, op = _problem(3)
    logscale = rng.uniform(-2.0, 2.0, gamma.shape)
    kind, binv, u, ci = co.grid_diagonal_preconditioner(gamma, noise, op, rank=8, logscale=logscale)
    v = rng.standard_normal((2, op.m, 3)) + 1j * rng.standard_normal((2, op.m, 3))
    got = co.apply_preconditioner(v, (kind, binv, u, ci))
    for b in range(2):
        order = np.argsort(-gamma[b], kind="stable")[:8]
        gt = gamma[b].copy()
        gt[order] = 0.0
        n = gamma.shape[1]
        w = np.exp(logscale[b]) / (n * (noise[b] + n * gt))
        bmat = np.linalg.inv((a * w) @ a.conj().T)
        cols = a[:, order] * np.sqrt(gamma[b, order])
        pmat = bmat + cols @ cols.conj().T
        np.testing.assert_allclose(got[b], np.linalg.solve(pmat, v[b]), rtol=1e-8, atol=1e-10)
        # Hermitian positive definite for any log-scale
        dense = co.apply_preconditioner(np.eye(op.m)[None].repeat(2, 0), (kind, binv, u, ci))[b]
        np.testing.assert_allclose(dense, dense.conj().T, atol=1e-10)
        assert np.linalg.eigvalsh(0.5 * (dense + dense.conj().T)).min() > 0


def test_grid_diagonal_preconditioner_is_q_for_at_most_32_atoms():
    """With at most 32 nonzero variances gamma_T = 0, B^-1 = A A^H / (N noise)
    = I / noise for any M < N, and B + F F^H = Q: PCG takes one step."""
    grid, samples = (16, 12), (9, 7)
    rng = np.random.default_rng(5)
    n = 16 * 12
    gamma = np.full((1, n), 0.0)
    gamma[0, rng.choice(n, 20, replace=False)] = rng.gamma(1.0, 1.0, 20)
    noise = np.array([0.03])
    op = co.Operator(grid, samples)
    pre = co.grid_diagonal_preconditioner(gamma, noise, op)
    v = rng.standard_normal((1, op.m, 2)) + 1j * rng.standard_normal((1, op.m, 2))
    np.testing.assert_allclose(co.apply_preconditioner(op.q(v, gamma, noise), pre), v, atol=1e-9)


def test_grid_diagonal_pcg_converges_to_the_dense_solution_in_fewer_steps():
    """PCG with the grid-diagonal background solves the same system as with
    the low-rank one (dense solution), on a clustered 32 x 32 / 20 x 20
    scene state, in no more iterations."""
    from fbme_amd import scenes
    sc = scenes.ellipse_scene(1000, (32, 32), (20, 20), k=60, snr_db=10.0)
    y = sc.y.reshape(1, -1).astype(np.complex128)
    op = co.Operator((32, 32), (20, 20))
    gamma = (np.abs(sc.x.reshape(1, -1)) ** 2 + 1e-3)
    noise = np.array([np.mean(np.abs(y) ** 2) * 0.02])
    a = dense_prefix_matrix((32, 32), (20, 20))
    exact = np.linalg.solve(noise[0] * np.eye(op.m) + (a * gamma[0]) @ a.conj().T, y[0])
    its = {}
    for kind in ("low_rank", "grid_diagonal"):
        x, it = co.pcg(y[:, :, None], gamma, noise, op, x0=None, tol=1e-10, max_iter=500, check=1, rank=32,
                       precond=kind, eps=np.finfo(np.float64).eps, tiny=1e-300)
        np.testing.assert_allclose(x[0, :, 0], exact, rtol=1e-8, atol=1e-8 * np.abs(exact).max())
        its[kind] = it
    assert its["grid_diagonal"] <= its["low_rank"], its
