"""Native Fast/Stable plans (HIP) against the float64 oracle and the
reference's golden vectors; graph replay; batched and uneven batches.

Parity bar: relative L2 <= 1e-4 on x, identical support mask and active set
(up to atoms whose polish score ties within float32 rounding).
"""

import numpy as np
import pytest
import torch

from conftest import golden_path
from fbme_amd import scenes
from oracle import fbme_oracle as fo

pytestmark = pytest.mark.gpu

GOLDEN = ["plan_32x32_24x24", "plan_40x32_20x16", "plan_64x64_40x40",
          "plan_96x80_60x48"]


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300))


def _make(profile, data, grid, device):
    from torch_sbl_fastest_stable import make_fastest_plan, make_stable_plan
    fn = make_fastest_plan if profile == "fast" else make_stable_plan
    return fn(data, grid, polish_size=32, warmup=1)


@pytest.mark.parametrize("case", GOLDEN)
@pytest.mark.parametrize("profile", ["fast", "stable"])
def test_plan_vs_golden_and_oracle(device, case, profile):
    print("case", case, profile, flush=True)
    g = np.load(golden_path(case + ".npz"))
    grid = tuple(int(v) for v in g["grid"])
    data = torch.from_numpy(g["y"]).to(device)
    plan = _make(profile, data, grid, device)
    res = plan.replay(synchronize=True)
    x = res.x.cpu().numpy().reshape(-1)
    orc = fo.fbme_plan(g["y"], grid, profile)
    xr = g[profile + "_x"].reshape(-1)
    assert _rel(x, orc["x"]) < 1e-4
    assert _rel(x, xr) < 1e-4
    assert set(np.flatnonzero(x)) == set(np.flatnonzero(orc["x"]))
    assert np.array_equal(res.support_mask.cpu().numpy().reshape(-1),
                          g[profile + "_support_mask"].reshape(-1))
    assert abs(float(res.noise_variance) / float(orc["noise_variance"]) - 1) < 1e-4
    assert abs(float(res.effective_dof) - float(orc["effective_dof"])) < 1e-3
    assert _rel(res.posterior_variance.cpu().numpy().reshape(-1),
                orc["posterior_variance"]) < 1e-4
    np.testing.assert_allclose(res.evidence_trace.cpu().numpy().reshape(-1),
                               orc["evidence_trace"], rtol=1e-5, atol=1e-2)
    assert res.iterations == int(g[profile + "_iterations"])
    if profile == "stable":
        assert int(res.support_exchanges_accepted) == int(orc["support_exchanges_accepted"])
        assert abs(float(res.fixed_budget_gamma_residual)
                   - float(orc["fixed_budget_gamma_residual"])) < 1e-4
        assert bool(res.fixed_budget_monotone) == bool(orc["fixed_budget_monotone"])


def test_uamp_sweep_state_vs_oracle(device):
    """Pre-polish UAMP state (final E-step) against the oracle sweep."""
    from fbme_amd import _lib
    from fbme_amd.plan import NativePlan
    sc = scenes.ellipse_scene(2000, (256, 256), (160, 160), k=300, snr_db=10.0)
    opt = _lib.default_options(False)
    plan = NativePlan([(256, 256, 160, 160)], opt, device)
    plan.store_uamp = True
    plan.run(torch.from_numpy(sc.y.reshape(-1)).to(device))
    torch.cuda.synchronize()
    n = 65536
    mean = plan.debug_tensor(1, torch.complex64, n).cpu().numpy()
    var = plan.debug_tensor(2, torch.float32, n).cpu().numpy()
    gam = plan.debug_tensor(0, torch.float32, n).cpu().numpy()
    sca = plan.debug_tensor(4, torch.float64, 8).cpu().numpy()
    y = sc.y.reshape(1, -1).astype(np.complex128)
    sw = fo.uamp_sweep(y, (256, 256), (160, 160), fo.PlanConfig.fast())
    assert _rel(mean, sw["mean"][0]) < 1e-4
    assert _rel(var, sw["variance"][0]) < 1e-4
    assert _rel(gam, sw["gamma"][0]) < 1e-4
    assert abs(sca[2] / sw["noise"][0] - 1) < 1e-5
    assert abs(sca[0] / sw["power"][0] - 1) < 1e-6


@pytest.mark.parametrize("profile", ["fast", "stable"])
def test_full_size_scene_vs_oracle(device, profile):
    """256^2 / 160^2, K=300 ellipse, 10 dB (the bench workload)."""
    sc = scenes.ellipse_scene(1000, (256, 256), (160, 160), k=300, snr_db=10.0)
    data = torch.from_numpy(sc.y).to(device)
    plan = _make(profile, data, (256, 256), device)
    res = plan.replay(synchronize=True)
    x = res.x.cpu().numpy().reshape(-1)
    orc = fo.fbme_plan(sc.y, (256, 256), profile)
    assert _rel(x, orc["x"]) < 1e-4
    assert np.array_equal(res.support_mask.cpu().numpy().reshape(-1),
                          orc["support_mask"])
    assert abs(float(res.noise_variance) / float(orc["noise_variance"]) - 1) < 1e-4


@pytest.mark.parametrize("snr_db", [0.0, 10.0, 20.0])
def test_stable_512_nonsquare_prefix_vs_oracle(device, snr_db):
    """512^2 grid, 384 x 256 prefix, K=1200 (six clustered ellipses plus 10 %
    isolated points), Stable plan, SNR swept over {0, 10, 20} dB."""
    sc = scenes.clustered_scene(1000, (512, 512), (384, 256), k=1200, snr_db=snr_db)
    data = torch.from_numpy(sc.y).to(device)
    print("512x512/384x256 stable snr", snr_db, flush=True)
    plan = _make("stable", data, (512, 512), device)
    res = plan.replay(synchronize=True)
    x = res.x.cpu().numpy().reshape(-1)
    orc = fo.fbme_plan(sc.y, (512, 512), "stable")
    assert _rel(x, orc["x"]) < 1e-4
    assert np.array_equal(res.support_mask.cpu().numpy().reshape(-1), orc["support_mask"])
    assert abs(float(res.noise_variance) / float(orc["noise_variance"]) - 1) < 1e-4
    assert int(plan.native_plan.out.status[0]) == 0


def test_graph_replay_new_data_matches_eager(device):
    from torch_sbl import SBLConfig, fast_sbl
    a = scenes.uniform_scene(3001, (64, 64), (40, 40), k=15, snr_db=15.0)
    b = scenes.uniform_scene(3002, (64, 64), (40, 40), k=15, snr_db=15.0)
    plan = _make("stable", torch.from_numpy(a.y).to(device), (64, 64), device)
    cfg = SBLConfig.realtime_hybrid_uamp(device=device, uamp_polish_size=32)
    for sc in (b, a, b):
        yb = torch.from_numpy(sc.y).to(device)
        res = plan.replay(yb, synchronize=True)
        eager = fast_sbl(yb, (64, 64), config=cfg)
        assert torch.equal(res.x, eager.x)
        assert torch.equal(res.noise_variance, eager.noise_variance)


def test_batched_equals_single(device):
    sc = [scenes.uniform_scene(4000 + i, (64, 64), (40, 40), k=12, snr_db=12.0)
          for i in range(5)]
    data = torch.from_numpy(np.stack([s.y for s in sc])).to(device)
    plan = _make("fast", data, (64, 64), device)
    res = plan.replay(synchronize=True)
    for i, s in enumerate(sc):
        single = _make("fast", torch.from_numpy(s.y).to(device), (64, 64), device)
        r1 = single.replay(synchronize=True)
        assert torch.equal(res.x[i], r1.x)
        assert torch.equal(res.noise_variance[i], r1.noise_variance)


def test_uneven_batch_equals_single(device):
    from fbme_amd import _lib
    from fbme_amd.plan import NativePlan
    batch = scenes.uneven_batch(77, count=6)
    shapes = [(s.grid[0], s.grid[1], s.samples[0], s.samples[1]) for s in batch]
    ys = torch.from_numpy(np.concatenate([s.y.reshape(-1) for s in batch])).to(device)
    opt = _lib.default_options(True)
    plan = NativePlan(shapes, opt, device)
    out = plan.run(ys)
    torch.cuda.synchronize()
    for i, s in enumerate(batch):
        go, so = plan.scene_slices()[i]
        n = s.grid[0] * s.grid[1]
        single = NativePlan([shapes[i]], opt, device)
        o1 = single.run(torch.from_numpy(s.y.reshape(-1)).to(device))
        torch.cuda.synchronize()
        assert torch.equal(out.x[go:go + n], o1.x)
        assert torch.equal(out.noise[i], o1.noise[0])
        orc = fo.fbme_plan(s.y, s.grid, "stable")
        assert _rel(o1.x.cpu().numpy(), orc["x"]) < 1e-4


MIXED = [(256, 256, 160, 160), (128, 256, 64, 200), (256, 128, 180, 100), (256, 256, 140, 256)]


@pytest.mark.parametrize("profile", ["fast", "stable"])
def test_register_fft_passes_match_lds_passes(device, profile):
    """The register 16 x 16 passes (every 256-point axis) against the LDS
    Stockham passes on the same batch, and both against the oracle.  The
    batch mixes scenes where only the rows, only the columns, or both axes
    have length 256."""
    from fbme_amd import _lib
    from fbme_amd.plan import NativePlan
    batch = [scenes.ellipse_scene(5000 + i, shp[:2], shp[2:], k=60, snr_db=20.0)
             for i, shp in enumerate(MIXED)]
    ys = torch.from_numpy(np.concatenate([s.y.reshape(-1) for s in batch])).to(device)
    outs = {}
    for kernel in (1, 0):
        print("fft_kernel", kernel, profile, flush=True)
        opt = _lib.default_options(profile == "stable")
        opt.fft_kernel = kernel
        plan = NativePlan(MIXED, opt, device)
        o = plan.run(ys)
        torch.cuda.synchronize()
        outs[kernel] = (plan.scene_slices(), o.x.cpu().numpy(), o.noise.cpu().numpy(),
                        o.support_mask.cpu().numpy(), o.status.cpu().numpy())
    (sl, x1, n1, m1, s1), (_, x0, n0, m0, s0) = outs[1], outs[0]
    assert not s0.any() and not s1.any()
    for i, s in enumerate(batch):
        go = sl[i][0]
        n = s.grid[0] * s.grid[1]
        orc = fo.fbme_plan(s.y, s.grid, profile)
        print("scene", i, MIXED[i], "register vs oracle", _rel(x0[go:go + n], orc["x"]),
              "lds vs oracle", _rel(x1[go:go + n], orc["x"]), flush=True)
        assert _rel(x0[go:go + n], x1[go:go + n]) < 1e-4
        assert _rel(x0[go:go + n], orc["x"]) < 1e-4
        assert set(np.flatnonzero(x0[go:go + n])) == set(np.flatnonzero(orc["x"]))
        assert np.array_equal(m0[go:go + n], orc["support_mask"])
        assert abs(n0[i] / n1[i] - 1) < 1e-4


def test_register_fft_is_selected_and_required(device):
    from fbme_amd import _lib
    from fbme_amd.plan import NativePlan
    opt = _lib.default_options(False)
    opt.fft_kernel = 2
    NativePlan([(256, 256, 160, 160)], opt, device)
    with pytest.raises(ValueError):
        NativePlan([(128, 256, 64, 160)], opt, device)
