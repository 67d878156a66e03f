"""Observation masks (sbl_2d / fast_sbl(..., observation_mask=)) on the
general solver (mode='cg') and the native UAMP plans, against the float64
oracles and the reference's golden runs.

The GPU path keeps full-length vectors with zeros on the unobserved samples
(the reference gathers the observed rows, operators.py:57-75 and 473-496).
Pieces: the mask spectrum against a direct float64 sum (and, for a full
mask, against the Dirichlet tables it generalises), the masked matvec and
preconditioner against the oracle's row-restricted operator, the masked
operator's forward / adjoint, and the whole solver against the reference's
golden runs (tests/golden/masked_*.npz, written by
tests/golden/make_golden.py).  Tolerances: float32 matvec relative L2
2e-5, preconditioner 1e-4; whole solves relative L2 1e-4 on x with
identical iteration counts and per-EM-iteration PCG counts and noise (the
achieved figures are stated per test).
"""

import numpy as np
import pytest
import torch

from conftest import assert_history_matches, golden_path
from oracle import cg_oracle as co
from oracle.fbme_oracle import F32_EPS, F32_TINY

pytestmark = pytest.mark.gpu

SHAPES = [((32, 32), (24, 24)), ((40, 30), (20, 15)), ((64, 48), (45, 40)), ((256, 256), (160, 160))]


def _mask(seed, samples, keep=0.6):
    return np.random.default_rng(seed).random(samples) < keep


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def scenes_mod():
    from fbme_amd import scenes
    return scenes


def _solver(device, grid, samples, mask, batch=2, seed=0):
    from fbme_amd import scenes
    from torch_sbl import SBLConfig, pcg
    ys = [np.where(mask, scenes.uniform_scene(seed + i, grid, samples, k=10, snr_db=20.0).y, 0)
          for i in range(batch)]
    y = torch.from_numpy(np.stack(ys).astype(np.complex64)).to(device)
    s = pcg._Solver(y, grid, SBLConfig(mode="cg", device=str(device)), None, torch.from_numpy(mask))
    s.init()
    torch.cuda.synchronize()
    return s


@pytest.mark.parametrize("grid,samples", [((40, 30), (20, 15)), ((256, 256), (160, 160))])
def test_mask_spectrum_vs_direct_sum(device, grid, samples):
    from fbme_amd import _lib
    from torch_sbl import pcg
    print("shape", grid, samples, flush=True)
    lib = _lib.load()
    n1, n2 = grid
    m1, m2 = samples
    stream = torch.cuda.current_stream(device).cuda_stream
    spec = torch.empty(n1 * n2, dtype=torch.complex128, device=device)
    work = torch.empty(m1 * n2, dtype=torch.complex128, device=device)
    mask = _mask(1, samples)
    mt = torch.from_numpy(mask.astype(np.uint8)).to(device)
    _lib.check(lib.fbme_mask_spectrum(stream, n1, n2, m1, m2, mt.data_ptr(), spec.data_ptr(), work.data_ptr()),
               "mask_spectrum")
    got = spec.cpu().numpy().reshape(n1, n2)
    e1 = np.exp(2j * np.pi * np.outer(np.arange(m1), np.arange(n1)) / n1)   # [m1, d1]
    e2 = np.exp(2j * np.pi * np.outer(np.arange(m2), np.arange(n2)) / n2)   # [m2, d2]
    want = e1.T @ mask.astype(np.float64) @ e2
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9 * mask.sum())
    # a full mask gives the product of the Dirichlet tables
    full = torch.ones(m1 * m2, dtype=torch.uint8, device=device)
    _lib.check(lib.fbme_mask_spectrum(stream, n1, n2, m1, m2, full.data_ptr(), spec.data_ptr(), work.data_ptr()),
               "mask_spectrum")
    t = pcg._tables(grid, samples, device)
    d1 = t.dir1.cpu().numpy().view(np.complex128)
    d2 = t.dir2.cpu().numpy().view(np.complex128)
    np.testing.assert_allclose(spec.cpu().numpy().reshape(n1, n2), np.outer(d1, d2), rtol=0,
                               atol=1e-9 * m1 * m2)


@pytest.mark.parametrize("grid,samples", SHAPES)
def test_masked_matvec_and_preconditioner_vs_oracle(device, grid, samples):
    print("shape", grid, samples, flush=True)
    mask = _mask(2, samples)
    s = _solver(device, grid, samples, mask, seed=30)
    assert s.obs is not None and not s.fused
    s.setup_operators()
    torch.cuda.synchronize()
    assert int(s.status.sum()) == 0
    keep = mask.reshape(-1)
    op = co.Operator(grid, samples, mask)
    gamma = s.gamma.cpu().numpy().astype(np.float64)
    noise = s.noise.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(6)
    d = (rng.standard_normal((s.b, s.r, s.m)) + 1j * rng.standard_normal((s.b, s.r, s.m))) * keep
    dt = torch.from_numpy(d.astype(np.complex64)).to(device)
    out = torch.empty_like(dt)
    dot = torch.empty(s.lanes, dtype=torch.float64, device=device)
    s.matvec(dt, out, dot=dot)
    got = out.cpu().numpy()
    assert not np.any(got[:, :, ~keep])
    ref = np.moveaxis(op.q(np.moveaxis(d[:, :, keep], 1, 2), gamma, noise), 2, 1)
    assert _rel(got[:, :, keep], ref) < 2e-5
    np.testing.assert_allclose(dot.cpu().numpy(), np.sum(np.real(np.conj(d) * got), -1).reshape(-1), rtol=1e-5)
    pre = co.low_rank_preconditioner(gamma, noise, op, 32, F32_EPS, F32_TINY)
    np.testing.assert_allclose(s.bg.cpu().numpy(), pre[0], rtol=1e-5)
    s.precondition(dt, out, dot=dot)
    got = out.cpu().numpy()
    assert not np.any(got[:, :, ~keep])
    ref = np.moveaxis(co.apply_preconditioner(np.moveaxis(d[:, :, keep], 1, 2), pre), 2, 1)
    assert _rel(got[:, :, keep], ref) < 1e-4
    np.testing.assert_allclose(dot.cpu().numpy(), np.sum(np.real(np.conj(d) * got), -1).reshape(-1), rtol=1e-5)


def test_masked_operator_forward_adjoint(device):
    from torch_sbl.operators import PrefixFourierOperator
    grid, samples = (40, 30), (20, 15)
    mask = _mask(3, samples)
    op = PrefixFourierOperator(grid, samples, device=device, observation_mask=torch.from_numpy(mask))
    ref = co.Operator(grid, samples, mask)
    assert op.n_samples == ref.m and not op.is_complete
    rng = np.random.default_rng(7)
    x = (rng.standard_normal((2, op.n_grid)) + 1j * rng.standard_normal((2, op.n_grid)))
    v = (rng.standard_normal((2, op.n_samples)) + 1j * rng.standard_normal((2, op.n_samples)))
    fx = op.forward_flat(torch.from_numpy(x.astype(np.complex64)).to(device)).cpu().numpy()
    av = op.adjoint_flat(torch.from_numpy(v.astype(np.complex64)).to(device)).cpu().numpy()
    assert _rel(fx, ref.fwd(x[:, :, None])[:, :, 0]) < 2e-6
    assert _rel(av, ref.adj(v[:, :, None])[:, :, 0]) < 2e-6
    idx = torch.tensor([[0, 7, 599, 1199]], device=device)
    cols = op.fourier_columns(idx).cpu().numpy()
    assert _rel(cols, ref.columns(idx.cpu().numpy())) < 1e-6


@pytest.mark.parametrize("name,grid,kw", [
    # achieved: relative L2 2.4e-6 on x, 76/76 iterations (active set at 30)
    ("masked_32x32_24x24", (32, 32), {}),
    # achieved: relative L2 4.3e-6 on x, 30/30 iterations
    ("masked_64x64_40x40_b2", (64, 64), dict(hutchinson_probes=16, auto_probe_scaling=False,
                                             cg_tolerance=1e-4, max_iter=30)),
])
def test_masked_sbl_2d_matches_reference_golden(device, name, grid, kw):
    """Against the reference's masked golden runs: identical iteration and
    per-EM-iteration PCG counts, noise within 1e-4 at every EM iteration,
    relative L2 1e-4 on x."""
    from fbme_amd import scenes
    from torch_sbl import SBLConfig, sbl_2d
    print("golden", name, flush=True)
    g = np.load(golden_path(name + ".npz"))
    y = torch.from_numpy(g["y"]).to(device)
    res = sbl_2d(y, grid, config=SBLConfig(mode="cg", device=str(device), **kw),
                 observation_mask=torch.from_numpy(g["mask"]))
    assert res.mode == "cg"
    assert res.iterations == int(g["cg_iterations"])
    assert res.active_set_size == int(g["cg_active_set_size"])
    dev = assert_history_matches(res, g)
    x = res.x.cpu().numpy()
    ref = g["cg_x"].reshape(x.shape)
    truth = g["x_true"].reshape(x.shape)
    xs, rs, ts = x.reshape(-1, *grid), ref.reshape(-1, *grid), truth.reshape(-1, *grid)
    for i in range(xs.shape[0]):
        print(name, i, "rel", _rel(xs[i], rs[i]), "noise history", dev, flush=True)
        assert _rel(xs[i], rs[i]) < 1e-4
        assert abs(scenes.nmse_db(xs[i], ts[i]) - scenes.nmse_db(rs[i], ts[i])) < 0.01
    np.testing.assert_allclose(res.noise_variance.cpu().numpy(), g["cg_noise_variance"], rtol=1e-4)
    np.testing.assert_allclose(res.effective_dof.cpu().numpy(), g["cg_effective_dof"], rtol=1e-4)


def test_masked_256_grid_vs_oracle(device):
    """256 x 256 grid, 160 x 160 prefix with 60% observed: three EM
    iterations of the masked solver against the float64 oracle (the grid
    matvec and FFT preconditioner with the mask applied to their outputs)."""
    from fbme_amd import scenes
    from torch_sbl import pcg, SBLConfig
    grid, samples = (256, 256), (160, 160)
    mask = _mask(4, samples)
    sc = scenes.ellipse_scene(1300, grid, samples, k=300, snr_db=10.0)
    y = np.where(mask, sc.y, 0).astype(np.complex64)
    kw = dict(max_iter=3, min_iter=1, tol=0.0, hutchinson_probes=16, auto_probe_scaling=False, cg_tolerance=1e-4)
    res = pcg.sbl_cg(torch.from_numpy(y[None]).to(device), grid, SBLConfig(mode="cg", device=str(device), **kw),
                     batch_shape=(1,), observation_mask=torch.from_numpy(mask))
    ref = co.sbl_cg(y[None], grid, co.CGConfig(**{k: v for k, v in kw.items() if k not in ("min_iter", "tol")}),
                    mask=mask)
    x = res.x.cpu().numpy().reshape(-1)
    xr = ref["x"].reshape(-1)
    print("masked 256:", _rel(x, xr), scenes.nmse_db(x, sc.x.reshape(-1)), scenes.nmse_db(xr, sc.x.reshape(-1)),
          "noise", float(res.noise_variance[0]) / float(ref["noise_variance"][0]) - 1,
          "pcg", res.history["cg_iterations"], ref["cg_iterations"], flush=True)
    # achieved: relative L2 7.4e-7, noise within 7e-7, PCG counts 8/8/16 as the oracle's
    assert _rel(x, xr) < 1e-4
    assert abs(scenes.nmse_db(x, sc.x.reshape(-1)) - scenes.nmse_db(xr, sc.x.reshape(-1))) < 0.01
    assert float(res.noise_variance[0]) == pytest.approx(float(ref["noise_variance"][0]), rel=1e-4)
    assert [int(v) for v in res.history["cg_iterations"]] == [int(v) for v in ref["cg_iterations"]]


def test_full_mask_is_unmasked_and_validation(device):
    from torch_sbl import SBLConfig, sbl_2d
    g = np.load(golden_path("general_32x32_24x24.npz"))
    y = torch.from_numpy(g["y"]).to(device)
    cfg = SBLConfig(mode="cg", device=str(device), max_iter=4)
    a = sbl_2d(y, (32, 32), config=cfg)
    b = sbl_2d(y, (32, 32), config=cfg, observation_mask=torch.ones(24, 24, dtype=torch.bool))
    assert torch.equal(a.x, b.x)
    with pytest.raises(ValueError):
        sbl_2d(y, (32, 32), config=cfg, observation_mask=torch.ones(24, 23, dtype=torch.bool))
    with pytest.raises(ValueError):
        sbl_2d(y, (32, 32), config=cfg, observation_mask=torch.zeros(24, 24, dtype=torch.bool))


# ---------------------------------------------------------------- UAMP plans
def _plan_cfg(profile, device):
    from torch_sbl import SBLConfig
    if profile == "fast":
        return SBLConfig.paper_hybrid_uamp(
            device=str(device), uamp_polish_size=32, uamp_polish_iterations=2,
            uamp_polish_residual_exchange_size=0, uamp_polish_fixed_budget_diagnostics=False,
            uamp_polish_compute_truncation_diagnostic=False)
    return SBLConfig.realtime_hybrid_uamp(
        device=str(device), uamp_polish_size=32, uamp_polish_iterations=3,
        uamp_polish_residual_exchange_size=12, uamp_polish_fixed_budget_diagnostics=True,
        uamp_polish_compute_truncation_diagnostic=False)


@pytest.mark.parametrize("case", ["masked_plan_64x64_40x40", "masked_plan_96x80_60x48"])
@pytest.mark.parametrize("profile", ["fast", "stable"])
def test_masked_plan_vs_golden_and_oracle(device, case, profile):
    """fast_sbl(..., observation_mask=) on the native plan against the
    reference's masked golden run and the float64 oracle, to the unmasked
    plans' bar (relative L2 1e-4, identical supports and support mask)."""
    from oracle import fbme_oracle as fo
    from torch_sbl import fast_sbl
    print("case", case, profile, flush=True)
    g = np.load(golden_path(case + ".npz"))
    grid = tuple(int(v) for v in g["grid"])
    res = fast_sbl(torch.from_numpy(g["y"]).to(device), grid, config=_plan_cfg(profile, device),
                   observation_mask=torch.from_numpy(g["mask"]))
    x = res.x.cpu().numpy().reshape(-1)
    orc = fo.fbme_plan(g["y"], grid, profile, mask=g["mask"])
    xr = g[profile + "_x"].reshape(-1)
    assert _rel(x, orc["x"]) < 1e-4
    assert _rel(x, xr) < 1e-4
    assert set(np.flatnonzero(x)) == set(np.flatnonzero(orc["x"]))
    assert np.array_equal(res.support_mask.cpu().numpy().reshape(-1), g[profile + "_support_mask"].reshape(-1))
    assert abs(float(res.noise_variance) / float(orc["noise_variance"]) - 1) < 1e-4
    assert abs(float(res.effective_dof) - float(orc["effective_dof"])) < 1e-3
    assert _rel(res.posterior_variance.cpu().numpy().reshape(-1), orc["posterior_variance"]) < 1e-4
    np.testing.assert_allclose(res.evidence_trace.cpu().numpy().reshape(-1), orc["evidence_trace"],
                               rtol=1e-5, atol=1e-2)
    assert res.iterations == int(g[profile + "_iterations"])
    if profile == "stable":
        assert int(res.support_exchanges_accepted) == int(orc["support_exchanges_accepted"])
        assert abs(float(res.fixed_budget_gamma_residual) - float(orc["fixed_budget_gamma_residual"])) < 1e-4
        assert bool(res.fixed_budget_monotone) == bool(orc["fixed_budget_monotone"])


@pytest.mark.parametrize("profile", ["fast", "stable"])
def test_masked_plan_256_register_kernels_vs_oracle(device, profile):
    """256 x 256 grid, 160 x 160 prefix with 60% observed (the register
    16 x 16 FFT passes with the mask)."""
    from oracle import fbme_oracle as fo
    from torch_sbl import fast_sbl
    grid, samples = (256, 256), (160, 160)
    mask = _mask(9, samples)
    sc = scenes_mod().ellipse_scene(1400, grid, samples, k=300, snr_db=10.0)
    res = fast_sbl(torch.from_numpy(sc.y).to(device), grid, config=_plan_cfg(profile, device),
                   observation_mask=torch.from_numpy(mask))
    x = res.x.cpu().numpy().reshape(-1)
    orc = fo.fbme_plan(sc.y, grid, profile, mask=mask)
    print("masked 256 plan", profile, _rel(x, orc["x"]), res.iterations, flush=True)
    assert res.iterations == orc["iterations"]
    assert _rel(x, orc["x"]) < 1e-4
    assert np.array_equal(res.support_mask.cpu().numpy().reshape(-1), orc["support_mask"])
    assert abs(float(res.noise_variance) / float(orc["noise_variance"]) - 1) < 1e-4


def test_uneven_masked_plan_equals_single_plans(device):
    """Masked and unmasked scenes of different shapes in one graph-captured
    plan give the single-scene plans' results exactly."""
    from fbme_amd.batched import UnevenPlan
    from fbme_amd.plan import NativePlan
    sc = scenes_mod()
    shapes = [(64, 64, 40, 40), (96, 80, 60, 48), (256, 256, 160, 160), (64, 64, 40, 40), (64, 64, 40, 40)]
    # the last scene repeats the first scene's mask (one shared mask spectrum)
    masks = [_mask(10, (40, 40)), None, _mask(11, (160, 160)), _mask(12, (40, 40), keep=0.8), _mask(10, (40, 40))]
    ys = [sc.uniform_scene(5000 + i, s[:2], s[2:], k=15, snr_db=15.0).y.astype(np.complex64)
          for i, s in enumerate(shapes)]
    plan = UnevenPlan(shapes, device=device, profile="stable", masks=masks, warmup=1)
    plan.replay([torch.from_numpy(y).to(device) for y in ys], synchronize=True)
    opt = plan.options
    for i, (shape, mk, y) in enumerate(zip(shapes, masks, ys)):
        single = NativePlan([shape], opt, device, [mk])
        single.run(torch.from_numpy(y.reshape(-1)).to(device))
        torch.cuda.synchronize()
        got = plan.scene(i)
        assert torch.equal(got["x"].reshape(-1), single.out.x), i
        assert torch.equal(got["noise_variance"], single.out.noise[0]), i
    # unobserved data values are ignored
    ys[0] = np.where(masks[0], ys[0], 1e3).astype(np.complex64)
    before = plan.scene(0)["x"].clone()
    plan.replay([torch.from_numpy(y).to(device) for y in ys], synchronize=True)
    assert torch.equal(plan.scene(0)["x"], before)


def test_full_mask_plan_is_unmasked_plan(device):
    from torch_sbl import fast_sbl
    g = np.load(golden_path("plan_64x64_40x40.npz"))
    y = torch.from_numpy(g["y"]).to(device)
    a = fast_sbl(y, (64, 64), config=_plan_cfg("stable", device))
    b = fast_sbl(y, (64, 64), config=_plan_cfg("stable", device),
                 observation_mask=torch.ones(40, 40, dtype=torch.bool))
    assert torch.equal(a.x, b.x)
