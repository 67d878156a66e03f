"""Exact dense posterior (mode='dense', solver.py:2962-2990) on the HIP
kernels against the float64 dense oracle and the reference's dense EM golden
runs.

The GPU path forms Q from the Fourier lags of gamma and factors it in FP64
with the GPU's deterministic loading (solver.py:2898-2920) at the FP64 eps,
8 eps_f64 mean|diag|; the data enters in complex64, the prefix transforms of
the inverse factor's rows run in FP64.  Against the
float64 oracle the bar is 1e-4 relative L2 on the posterior mean, variance
and data diagonal of one step; against the reference's complex128
dense_em_sbl runs (30 and 200 EM iterations) 1e-5 on x, 1e-3 on the noise
variance (the runs end with M - dof = 0.2 in its denominator, which turns a
1e-6 relative difference in dof into 3e-4 in the noise) and 1e-5 on the
dof.
"""

import numpy as np
import pytest
import torch

from conftest import golden_path
from oracle import dense_em
from oracle import fbme_oracle as fo

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _dense_cfg(device, iters):
    from torch_sbl import SBLConfig
    # literature_reproduction/baselines.py:254-278
    return SBLConfig(mode="dense", max_iter=iters, min_iter=iters, tol=0.0, hyper_shape=1e-6, hyper_rate=1e-6,
                     noise_shape=1e-6, noise_rate=1e-6, initialization="uniform", damping=0.0,
                     noise_update="evidence", active_set=False, device=str(device), return_history=False,
                     eager_change_tracking=False)


@pytest.mark.parametrize("iters", [30, 200])
def test_dense_em_matches_reference_golden(device, iters):
    from torch_sbl import sbl_2d
    g = np.load(golden_path("general_32x32_24x24.npz"))
    res = sbl_2d(torch.from_numpy(g["y"]).to(device), (32, 32), config=_dense_cfg(device, iters))
    assert res.mode == "dense" and res.iterations == iters
    pre = f"dense{iters}_"
    x = res.x.cpu().numpy()
    print("dense", iters, _rel(x, g[pre + "x"]), flush=True)
    assert _rel(x, g[pre + "x"]) < 1e-5
    assert float(res.noise_variance) == pytest.approx(float(g[pre + "noise_variance"]), rel=1e-3)
    assert float(res.effective_dof) == pytest.approx(float(g[pre + "effective_dof"]), rel=1e-5)


@pytest.mark.parametrize("grid,samples,masked", [((32, 32), (24, 24), False), ((40, 30), (20, 15), True),
                                                 ((64, 64), (40, 40), False)])
def test_dense_posterior_step_vs_oracle(device, grid, samples, masked):
    """One posterior evaluation at random variances against the float64
    dense posterior on the (row-restricted) dense matrix."""
    from torch_sbl import SBLConfig, pcg
    from fbme_amd import scenes
    print("dense step", grid, samples, masked, flush=True)
    rng = np.random.default_rng(3)
    mask = rng.random(samples) < 0.7 if masked else None
    sc = scenes.uniform_scene(70, grid, samples, k=12, snr_db=20.0)
    y = sc.y if mask is None else np.where(mask, sc.y, 0)
    yt = torch.from_numpy(y[None].astype(np.complex64)).to(device)
    s = pcg._Solver(yt, grid, SBLConfig(mode="dense", device=str(device)), None,
                    None if mask is None else torch.from_numpy(mask), dense=True)
    s.init()
    n = grid[0] * grid[1]
    gamma = rng.gamma(0.5, 1.0, n) * 0.01
    s.gamma.copy_(torch.from_numpy(gamma[None].astype(np.float32)))
    s.noise.fill_(0.05)
    s.posterior(0.0)
    torch.cuda.synchronize()
    assert int(s.status.sum()) == 0
    a = fo.dense_prefix_matrix(grid, samples)
    yv = y.reshape(-1).astype(np.complex128)
    if mask is not None:
        keep = mask.reshape(-1)
        a, yv = a[keep], yv[keep]
    mean, var, diag = dense_em.dense_posterior(a, yv, gamma.astype(np.float32).astype(np.float64), 0.05)
    assert _rel(s.mean.cpu().numpy()[0], mean) < 1e-4
    assert _rel(s.var.cpu().numpy()[0], var) < 1e-4
    assert _rel(s.dd.cpu().numpy()[0], diag) < 1e-4


def test_dense_batch_matches_single_scenes(device):
    """Three scenes in one launch (scene index on the grid's y / z axes of
    every dense kernel) give each scene's single-launch posterior, and the
    float64 dense posterior.  40x30 / 20x15: M = 300, ten block rows, so the
    inverse's partial sums and the row chunks of the data diagonal are
    exercised."""
    from torch_sbl import SBLConfig, pcg
    from fbme_amd import scenes
    grid, samples, b = (40, 30), (20, 15), 3
    n = grid[0] * grid[1]
    rng = np.random.default_rng(5)
    ys = np.stack([scenes.uniform_scene(80 + i, grid, samples, k=10, snr_db=20.0).y for i in range(b)])
    gammas = (rng.gamma(0.5, 1.0, (b, n)) * 0.01).astype(np.float32)
    noises = np.array([0.03, 0.05, 0.08], np.float32)
    cfg = SBLConfig(mode="dense", device=str(device))

    def run(y, g, nu):
        s = pcg._Solver(torch.from_numpy(y.astype(np.complex64)).to(device), grid, cfg, None, None, dense=True)
        s.init()
        s.gamma.copy_(torch.from_numpy(g))
        s.noise.copy_(torch.from_numpy(nu))
        s.posterior(0.0)
        torch.cuda.synchronize()
        assert int(s.status.sum()) == 0
        return s.mean.cpu().numpy(), s.var.cpu().numpy(), s.dd.cpu().numpy()

    batch = run(ys, gammas, noises)
    for i in range(b):
        single = run(ys[i:i + 1], gammas[i:i + 1], noises[i:i + 1])
        for got, want in zip(batch, single):
            assert _rel(got[i], want[0]) < 1e-6
    a = fo.dense_prefix_matrix(grid, samples)
    mean, var, diag = dense_em.dense_posterior(a, ys[1].reshape(-1).astype(np.complex128),
                                               gammas[1].astype(np.float64), float(noises[1]))
    assert _rel(batch[0][1], mean) < 1e-4
    assert _rel(batch[1][1], var) < 1e-4
    assert _rel(batch[2][1], diag) < 1e-4


def test_auto_mode_small_problem_runs_dense(device):
    """Default sbl_2d on the GPU with M <= dense_threshold takes the exact
    dense posterior, as the reference does (solver.py:1025-1036)."""
    from fbme_amd import scenes
    from torch_sbl import SBLConfig, sbl_2d
    sc = scenes.uniform_scene(71, (32, 32), (24, 24), k=12, snr_db=20.0)
    res = sbl_2d(torch.from_numpy(sc.y).to(device), (32, 32), config=SBLConfig(device=str(device), max_iter=20))
    assert res.mode == "dense"
    assert scenes.nmse_db(res.x.cpu().numpy().reshape(-1), sc.x.reshape(-1)) < -15.0


def test_dense_posterior_rejects_bad_arguments(device):
    """C-ABI error paths: a workspace smaller than the layout is refused
    before any launch, and an observed-sample index outside [0, M1 M2) is
    left out by the kernels and reported as status 2."""
    import ctypes
    from fbme_amd import _lib
    lib = _lib.load()
    b, n1, n2, m1, m2 = 1, 16, 16, 8, 8
    n, mf, m = n1 * n2, m1 * m2, 30
    nbytes = int(lib.fbme_dense_workspace_bytes(b, n1, n2, m1, m2, m))
    assert nbytes > 0
    work = torch.empty(nbytes, dtype=torch.uint8, device=device)
    y = torch.randn(b, mf, dtype=torch.complex64, device=device)
    gamma = torch.full((b, n), 0.1, dtype=torch.float32, device=device)
    noise = torch.full((b,), 0.01, dtype=torch.float32, device=device)
    mean = torch.empty(b, n, dtype=torch.complex64, device=device)
    var = torch.empty(b, n, dtype=torch.float32, device=device)
    dd = torch.empty_like(var)
    part = torch.empty(b * ((n + 255) // 256), dtype=torch.float64, device=device)
    rel = torch.empty(b, dtype=torch.float64, device=device)
    status = torch.zeros(b, dtype=torch.int32, device=device)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)

    def call(obs, size):
        return lib.fbme_dense_posterior(stream, b, n1, n2, m1, m2, m, p(obs), p(y), p(gamma), p(noise),
                                        8.0 * 2.0 ** -52, 1e-38, p(work), size, p(mean), p(var), p(dd), p(part),
                                        p(rel), p(status))

    good = torch.arange(0, 2 * m, 2, dtype=torch.int32, device=device)   # 0, 2, .., 58 < 64
    assert call(good, nbytes - 1) == _lib.ERR_ARG
    assert call(good, nbytes) == 0
    torch.cuda.synchronize()
    assert int(status.sum()) == 0
    assert bool(torch.isfinite(var).all())
    bad = good.clone()
    bad[5] = mf + 3
    assert call(bad, nbytes) == 0
    torch.cuda.synchronize()
    assert int(status[0]) == 2
