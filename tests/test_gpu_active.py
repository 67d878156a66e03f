"""The general solver's active-set switch beyond 32 atoms
(_select_active_indices, solver.py:3184-3216; switch 1316-1341; reduced
posterior 3591-3779, 3897-3924) on the HIP kernels.

* fbme_topk (k <= 256) against numpy's stable argsort, with ties;
* fbme_reduced_posterior_large (FP64 tiled factorisation, K up to 256)
  against the float64 oracle's reduced posterior and, at K = 32, against
  the one-wavefront kernel of the K <= 32 path;
* whole solves against the reference's golden runs
  (tests/golden/active_*.npz): switches to 58, 161 and 37 atoms, a
  maximum of 16 atoms (never switches) and a minimum of 48 atoms, with
  identical switch iteration, active-set size and iteration count, per
  EM iteration noise within 1e-4 and identical PCG counts (at most one EM
  iteration may stop its PCG one convergence check apart, a tolerance
  decision float32 rounding makes), and relative L2 1e-4 on x (achieved
  figures printed).
"""

import ctypes

import numpy as np
import pytest
import torch

from conftest import assert_history_matches, golden_path
from oracle import cg_oracle as co
from oracle.fbme_oracle import F32_EPS, F32_TINY

pytestmark = pytest.mark.gpu

ACTIVE = ["active_64x64_40x40_k60", "active_48x48_32x32_k50", "active_64x64_40x40_k45",
          "active_64x64_40x40_k60_max16", "active_64x64_40x40_k45_min48"]


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n,k", [(4096, 256), (65536, 256), (262144, 100), (5000, 37)])
def test_topk_general_vs_stable_argsort(device, n, k):
    from fbme_amd import _lib
    lib = _lib.load()
    b = 2
    rng = np.random.default_rng(n + k)
    x = (rng.integers(0, 200, size=(b, n)) / 200.0).astype(np.float32)   # many exact ties
    xt = torch.from_numpy(x).to(device)
    cand = torch.empty(int(lib.fbme_topk_scratch_bytes(b, n, k)) // 8, dtype=torch.int64, device=device)
    idx = torch.empty((b, k), dtype=torch.int32, device=device)
    val = torch.empty((b, k), dtype=torch.float32, device=device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    _lib.check(lib.fbme_topk(stream, b, n, k, _p(xt), _p(cand), _p(idx), _p(val)), "topk")
    torch.cuda.synchronize()
    want = np.argsort(-x, axis=1, kind="stable")[:, :k]
    np.testing.assert_array_equal(idx.cpu().numpy(), want)
    np.testing.assert_array_equal(val.cpu().numpy(), np.take_along_axis(x, want, 1))


@pytest.mark.parametrize("k,masked", [(32, False), (100, False), (256, False), (120, True)])
def test_large_reduced_posterior_vs_oracle(device, k, masked):
    from fbme_amd import _lib, scenes
    from torch_sbl import SBLConfig, pcg
    grid, samples = (64, 64), (40, 40)
    sc = scenes.uniform_scene(1090, grid, samples, k=80, snr_db=20.0)
    mask = (np.random.default_rng(5).random(samples) < 0.7) if masked else None
    y = sc.y if mask is None else np.where(mask, sc.y, 0).astype(np.complex64)
    s = pcg._Solver(torch.from_numpy(y[None]).to(device), grid, SBLConfig(mode="cg", device=str(device)),
                    observation_mask=None if mask is None else torch.from_numpy(mask))
    s.init()
    # a spread of variances: the backprojection init plus a random factor
    g = s.gamma.cpu().numpy() * np.random.default_rng(6).uniform(0.5, 4.0, size=s.gamma.shape)
    s.gamma.copy_(torch.from_numpy(g.astype(np.float32)))
    s._chk(s.lib.fbme_grid_sums(s.stream, s.b, s.n, _p(s.gamma), _p(s.mean), _p(s.prev), _p(s.grid_part),
                                _p(s.sums)), "sums")
    sup = np.argsort(-g, axis=1, kind="stable")[:, :k].astype(np.int32)
    s.support = torch.from_numpy(sup).to(device)
    s.rp_work = torch.empty(int(s.lib.fbme_reduced_posterior_large_bytes(1, k)), dtype=torch.uint8, device=device)
    t = s.t
    gsum = s.sums[:, 0].contiguous()
    _lib.check(s.lib.fbme_reduced_posterior_large(
        s.stream, 1, k, *grid, _p(s.support), _p(s.gamma), _p(gsum), _p(s.noise), _p(s.bp), _p(t.dir1), _p(t.dir2),
        _p(s.mask_spec), 8.0, _p(s.rp_work), s.rp_work.numel(), _p(s.mean), _p(s.var), _p(s.dd), _p(s.status)),
        "large")
    torch.cuda.synchronize()
    assert int(s.status[0]) == 0
    mean, var, dd = s.mean.cpu().numpy()[0], s.var.cpu().numpy()[0], s.dd.cpu().numpy()[0]
    op = co.Operator(grid, samples, mask)
    yy = y.reshape(1, -1).astype(np.complex128)
    if mask is not None:
        yy = yy[:, mask.reshape(-1)]
    gg = s.gamma.cpu().numpy().astype(np.float64)
    noise = s.noise.cpu().numpy().astype(np.float64)
    om, ov, od = co.active_posterior(yy, gg, noise, sup.astype(np.int64), op, F32_EPS, F32_TINY)
    print("K", k, "masked", masked, "mean", _rel(mean, om[0]), "var", _rel(var, ov[0]), "dd", _rel(dd, od[0]),
          flush=True)
    assert _rel(mean, om[0]) < 1e-5
    assert _rel(var, ov[0]) < 1e-5
    assert _rel(dd[sup[0]], od[0][sup[0]]) < 1e-5
    if k == 32:
        # the K <= 32 one-wavefront kernel on the same support agrees
        m32, v32 = s.mean.clone(), s.var.clone()
        s.posterior(1e-5)
        torch.cuda.synchronize()
        assert _rel(s.mean.cpu().numpy()[0], m32.cpu().numpy()[0]) < 1e-6
        assert _rel(s.var.cpu().numpy()[0], v32.cpu().numpy()[0]) < 1e-6


@pytest.mark.parametrize("case", ACTIVE)
def test_active_set_sizes_match_reference(device, case):
    from torch_sbl import SBLConfig, sbl_2d
    print("case", case, flush=True)
    g = np.load(golden_path(case + ".npz"))
    over = {k[len("override_"):]: g[k].item() for k in g.files if k.startswith("override_")}
    grid = tuple(int(v) for v in g["grid"])
    res = sbl_2d(torch.from_numpy(g["y"]).to(device), grid,
                 config=SBLConfig(mode="cg", device=str(device), **over))
    np.testing.assert_array_equal(np.asarray(res.history["active_set_size"]),
                                  g["cg_history_active_set_size"])
    assert res.active_set_size == int(g["cg_active_set_size"])
    assert res.iterations == int(g["cg_iterations"])
    # the 37-atom scene and the two size-limit variants have one EM
    # iteration whose PCG stops one check apart from the reference
    dev = assert_history_matches(res, g, pcg_slack=1)
    x = res.x.cpu().numpy().reshape(-1)
    rel = _rel(x, g["cg_x"].reshape(-1))
    print(case, "active", res.active_set_size, "rel", rel, "noise history", dev, flush=True)
    assert rel < 1e-4
    assert float(res.noise_variance) == pytest.approx(float(g["cg_noise_variance"]), rel=1e-4)
