"""The grid-diagonal PCG preconditioner and its learned correction.

The preconditioner is the low-rank Woodbury form of the reference
(solver.py:5786-5848) with its scalar background replaced by
B^-1 = A diag(w) A^H, w = exp(clamp(o)) / (N (noise + N gamma_T)), o the
preconditioner network's per-cell output (0: the closed form).  Checks:

* the solver's preconditioner apply (fbme_gdiag_weights, fbme_lowrank_columns,
  fbme_gdiag_system, fbme_lowrank_apply_split and the B^-1 matvec passes)
  against the float64 oracle, on the BTTB passes (64x64 / 40x40) and the
  256-point grid passes (256x256 / 160x160), with and without a network
  output: relative L2 <= 2e-5 (float32 transforms);
* the network's input features and the backward pass from an external loss
  gradient (fbme_cnn_output_grad) against the MSE path of fbme_cnn_loss;
* whole solves at 256x256 / 160x160, K = 300 ellipse scenes, 10 dB, on
  held-out seeds: the grid-diagonal background, closed form and learned,
  takes fewer PCG iterations than the low-rank preconditioner for the same
  EM budget at NMSE within 0.1 dB of it; a stalled learned solve falls back
  to the plain solver's result exactly.
"""

import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GRID, SAMPLES, K, SNR = (256, 256), (160, 160), 300, 10.0
SEEDS = range(1430, 1436)


def _cfg(device, iters, **kw):
    from torch_sbl import SBLConfig
    return SBLConfig(mode="cg", device=str(device), max_iter=iters, min_iter=1, tol=0.0,
                     hutchinson_probes=16, auto_probe_scaling=False, cg_tolerance=1e-4, **kw)


def _state(device, grid, samples, seed, net=None):
    """A solver in grid-diagonal mode whose variances are a random positive
    field with distinct values and a few strong cells."""
    from fbme_amd import scenes
    from torch_sbl import pcg
    sc = scenes.ellipse_scene(seed, grid, samples, k=max(8, grid[0] * grid[1] // 200), snr_db=10.0)
    y = torch.from_numpy(sc.y[None].astype(np.complex64)).to(device)
    s = pcg._Solver(y, grid, _cfg(device, 1), precond="grid_diagonal", precond_net=net)
    s.init()
    rng = np.random.default_rng(seed)
    n = grid[0] * grid[1]
    g = rng.gamma(0.3, 1.0, n) * 1e-2
    g[rng.choice(n, 60, replace=False)] += rng.gamma(2.0, 1.0, 60)
    s.gamma.copy_(torch.from_numpy(g.astype(np.float32)[None]))
    s.noise.fill_(float(np.mean(g) * 3.0))
    s.grid_sums(with_mean=False)
    s.setup_operators()
    return s


@pytest.mark.parametrize("grid,samples", [((64, 64), (40, 40)), ((256, 256), (160, 160))])
@pytest.mark.parametrize("with_net", [False, True])
def test_preconditioner_matches_oracle(device, grid, samples, with_net):
    from oracle import cg_oracle as co
    from torch_sbl import learned
    net = learned.PrecondNet(device, seed=3) if with_net else None
    s = _state(device, grid, samples, 7, net)
    print(f"grid_diagonal apply {grid}/{samples} net={with_net}", flush=True)
    rng = np.random.default_rng(2)
    v = (rng.standard_normal((s.lanes, s.m)) + 1j * rng.standard_normal((s.lanes, s.m))).astype(np.complex64)
    vt = torch.from_numpy(v).to(device)
    out = torch.empty_like(vt)
    dot = torch.empty(s.lanes, dtype=torch.float64, device=device)
    s.precondition(vt, out, dot=dot)
    torch.cuda.synchronize()
    assert int(s.status.sum()) == 0
    gamma = s.gamma.double().cpu().numpy()
    noise = s.noise.double().cpu().numpy()
    logscale = s.gd_logscale.reshape(1, -1).double().cpu().numpy() if with_net else None
    if with_net:
        assert float(np.abs(logscale).max()) > 1e-3        # the network's output reaches w
    op = co.Operator(grid, samples)
    pre = co.grid_diagonal_preconditioner(gamma, noise, op, logscale=logscale)
    ref = co.apply_preconditioner(v.T[None].astype(np.complex128), pre)[0].T
    got = out.cpu().numpy()
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"  relative L2 {rel:.2e}", flush=True)
    assert rel < 2e-5
    rz = np.real(np.sum(np.conj(v) * got, axis=1))
    np.testing.assert_allclose(dot.cpu().numpy(), rz, rtol=1e-6)


def test_precond_features_and_external_gradient(device):
    from torch_sbl import learned
    net = learned.PrecondNet(device, seed=5)
    s = _state(device, (64, 64), (40, 40), 9, None)
    x = net.features(s.bp, s.gamma, s.bpe, s.sums[:, 0].contiguous(), s.noise, (64, 64))
    bp = s.bp.cpu().numpy().reshape(-1)
    g = s.gamma.double().cpu().numpy().reshape(-1)
    noise = float(s.noise[0])
    e = np.abs(bp) ** 2
    f0 = np.clip(np.log(e / (float(s.bpe[0]) / g.size) + 1e-6), -30, 30)
    f1 = np.clip(np.log(np.maximum(g, 1e-37) / (float(s.sums[0, 0]) / g.size)), -30, 30)
    f2 = np.minimum(np.log1p(g.size * g / noise) / 8, 30)
    xf = (x.cpu().numpy().astype(np.uint16).astype(np.uint32) << 16).view(np.float32).reshape(-1, 16)
    for c, ref in enumerate((f0, f1, f2)):
        np.testing.assert_allclose(xf[:, c], ref, rtol=1e-2, atol=2e-2)
    assert not xf[:, 3:].any()
    # backward from g = 2 (out - t) / count equals the MSE path
    t = torch.randn((1, 64, 64), device=device)
    net.loss_and_grads(x, t)
    ref_grads = net.grads.clone()
    out = net.forward(x)
    grad = (2.0 * (out - t) / out.numel()).contiguous()
    net.backward_from(x, grad)
    torch.cuda.synchronize()
    assert torch.equal(net.grads, ref_grads)


def _solve_all(device, y, iters, **kw):
    from torch_sbl import learned, pcg
    out = []
    for i in range(y.shape[0]):
        yt = torch.from_numpy(y[i:i + 1]).to(device)
        cfg = _cfg(device, iters, return_history=True)
        if "precond_net" in kw:
            out.append(learned.learned_sbl_2d(yt, GRID, None, cfg, precond_net=kw["precond_net"]))
        else:
            out.append(pcg.sbl_cg(yt, GRID, cfg, batch_shape=(1,), **kw))
    return out


@pytest.fixture(scope="module")
def held_out():
    from fbme_amd import scenes
    sc = [scenes.ellipse_scene(s, GRID, SAMPLES, k=K, snr_db=SNR) for s in SEEDS]
    return sc, np.stack([s.y for s in sc]).astype(np.complex64)


def _stats(rs, sc):
    from fbme_amd import scenes
    nm = [scenes.nmse_db(r.x.reshape(-1).cpu().numpy(), s.x.reshape(-1)) for r, s in zip(rs, sc)]
    its = [float(np.sum(r.history["cg_iterations"])) for r in rs]
    return float(np.mean(nm)), float(np.mean(its))


@pytest.mark.parametrize("iters", [3, 10])
def test_grid_diagonal_needs_fewer_pcg_iterations(device, held_out, iters):
    sc, y = held_out
    plain = _stats(_solve_all(device, y, iters), sc)
    gd = _stats(_solve_all(device, y, iters, precond="grid_diagonal"), sc)
    print(f"{iters} EM iterations: low-rank (NMSE dB, PCG its) {plain}, grid-diagonal {gd}", flush=True)
    assert abs(gd[0] - plain[0]) < 0.1
    assert gd[1] < 0.75 * plain[1]


@pytest.fixture(scope="module")
def pnet(device):
    from torch_sbl import learned
    assert os.path.exists(learned.PRECOND_WEIGHTS), "trained weights missing: " + learned.PRECOND_WEIGHTS
    return learned.PrecondNet.load(device)


@pytest.mark.parametrize("iters", [3, 10])
def test_learned_preconditioner_needs_fewer_pcg_iterations(device, held_out, pnet, iters):
    sc, y = held_out
    plain = _stats(_solve_all(device, y, iters), sc)
    rs = _solve_all(device, y, iters, precond_net=pnet)
    lp = _stats(rs, sc)
    print(f"{iters} EM iterations: low-rank {plain}, learned grid-diagonal {lp}", flush=True)
    assert all(r.history["learned_fallback"] == [0.0] for r in rs)
    assert abs(lp[0] - plain[0]) < 0.1
    assert lp[1] < 0.75 * plain[1]


def test_learned_preconditioner_fallback_returns_plain_solution(device, held_out, pnet):
    """A PCG budget of 2 iterations stalls the solve: the batch is solved
    again by the plain solver (low-rank preconditioner, plain start)."""
    from torch_sbl import learned, pcg
    _, y = held_out
    yt = torch.from_numpy(y[:2]).to(device)
    cfg = _cfg(device, 2, cg_max_iter=2)
    r = learned.learned_sbl_2d(yt, GRID, None, cfg, precond_net=pnet)
    p = pcg.sbl_cg(yt, GRID, cfg, batch_shape=(2,))
    assert r.history["learned_fallback"] == [1.0]
    assert torch.equal(r.x, p.x)
    assert torch.equal(r.noise_variance, p.noise_variance)
