"""Autotuned pass shapes change speed, not results: tuned uneven plans and
tiled BTTB matvecs against their default shapes, and the uneven graph plan
against single-scene plans."""

import numpy as np
import pytest
import torch

from fbme_amd import scenes

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_tuner_records_fastest_and_results_agree(device, tmp_path, monkeypatch):
    from fbme_amd import _lib, autotune
    from fbme_amd.batched import UnevenPlan
    from fbme_amd.plan import NativePlan
    monkeypatch.setenv("FBME_TUNING", str(tmp_path / "tuning.json"))
    batch = scenes.uneven_batch(11, count=8)
    shapes = [(s.grid[0], s.grid[1], s.samples[0], s.samples[1]) for s in batch]
    cands = [{"fft_kernel": 0, "row_lines": 4, "col_lines": 8}, {"fft_kernel": 1, "row_lines": 2, "col_lines": 4},
             {"fft_kernel": 0, "row_lines": 16, "col_lines": 16}]
    entry = autotune.tune_plan(shapes, False, device, reps=3, candidates=cands)
    print("tuned", entry, flush=True)
    assert autotune.lookup(autotune.device_arch(device), autotune.plan_signature(shapes, False)) == entry
    ys = [torch.from_numpy(s.y).to(device) for s in batch]
    tuned = UnevenPlan(shapes, device=device)
    assert tuned.options.row_lines == entry["row_lines"] and tuned.options.fft_kernel == entry["fft_kernel"]
    tuned.replay(ys, synchronize=True)
    base = NativePlan(shapes, _lib.default_options(False), device)
    packed = tuned.input.clone()
    out = base.run(packed)
    torch.cuda.synchronize()
    for i, s in enumerate(shapes):
        go = base.scene_slices()[i][0]
        n = s[0] * s[1]
        got = tuned.scene(i)["x"].reshape(-1).cpu().numpy()
        want = out.x[go:go + n].cpu().numpy()
        assert _rel(got, want) < 1e-4, i


def test_uneven_graph_plan_equals_single_scenes(device, monkeypatch, tmp_path):
    from fbme_amd import _lib
    from fbme_amd.batched import UnevenPlan
    from fbme_amd.plan import NativePlan
    monkeypatch.setenv("FBME_TUNING", str(tmp_path / "none.json"))
    batch = scenes.uneven_batch(5, count=6)
    shapes = [(s.grid[0], s.grid[1], s.samples[0], s.samples[1]) for s in batch]
    plan = UnevenPlan(shapes, device=device, profile="stable", options=_lib.default_options(True))
    ys = [torch.from_numpy(s.y).to(device) for s in batch]
    plan.replay(ys, synchronize=True)
    first = [plan.scene(i)["x"].clone() for i in range(len(batch))]
    plan.replay(synchronize=True)
    for i, s in enumerate(batch):
        assert torch.equal(plan.scene(i)["x"], first[i])
        single = NativePlan([shapes[i]], _lib.default_options(True), device)
        o = single.run(ys[i].reshape(-1))
        torch.cuda.synchronize()
        assert torch.equal(first[i].reshape(-1), o.x), i
        assert torch.equal(plan.scene(i)["noise_variance"], o.noise[0])


@pytest.mark.parametrize("grid,samples", [((256, 256), (160, 160)), ((64, 48), (45, 40))])
def test_tiled_toeplitz_matvec_is_bitwise_equal(device, grid, samples):
    import ctypes
    from fbme_amd import _lib
    from torch_sbl import SBLConfig, pcg
    ys = [scenes.uniform_scene(i, grid, samples, k=10, snr_db=20.0).y for i in range(2)]
    y = torch.from_numpy(np.stack(ys).astype(np.complex64)).to(device)
    s = pcg._Solver(y, grid, SBLConfig(mode="cg", device=str(device)))
    s.grid_op = False   # the BTTB passes
    s.init()
    s.setup_operators()
    lib = _lib.load()
    rmax, cmax = ctypes.c_int(), ctypes.c_int()
    _lib.check(lib.fbme_toeplitz_tiles(s.t.e1, s.t.e2, ctypes.byref(rmax), ctypes.byref(cmax)), "tiles")
    ref = torch.empty_like(s.rhs)
    s.tiles = (0, 0)
    s.matvec(s.rhs, ref)
    for rt, ct in [(1, 1), (2, cmax.value), (rmax.value, 1), (3, 5)]:
        out = torch.empty_like(s.rhs)
        s.tiles = (rt, ct)
        s.matvec(s.rhs, out)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (rt, ct)


@pytest.mark.parametrize("samples", [(160, 160), (96, 200), (256, 40)])
def test_grid_matvec_matches_bttb_and_oracle(device, samples):
    """Q d through the 256 x 256 grid (register FFT passes) against the BTTB
    passes and the float64 operator, on the solver's own state and lanes."""
    from oracle import cg_oracle as co
    from torch_sbl import SBLConfig, pcg
    grid = (256, 256)
    ys = [scenes.ellipse_scene(20 + i, grid, samples, k=100, snr_db=10.0).y for i in range(2)]
    y = torch.from_numpy(np.stack(ys).astype(np.complex64)).to(device)
    s = pcg._Solver(y, grid, SBLConfig(mode="cg", device=str(device), hutchinson_probes=4,
                                        auto_probe_scaling=False))
    assert s.grid_op
    s.init()
    s.setup_operators()
    s.spectrum_setup()
    got = torch.empty_like(s.rhs)
    s.matvec(s.rhs, got)
    s.grid_op = False
    ref = torch.empty_like(s.rhs)
    s.matvec(s.rhs, ref)
    torch.cuda.synchronize()
    g, r = got.cpu().numpy(), ref.cpu().numpy()
    assert _rel(g, r) < 2e-5
    op = co.Operator(grid, samples)
    d = np.moveaxis(s.rhs.cpu().numpy().astype(np.complex128), 1, 2)      # [B, M, R]
    q = op.q(d, s.gamma.cpu().numpy().astype(np.float64), s.noise.cpu().numpy().astype(np.float64))
    assert _rel(g, np.moveaxis(q, 2, 1)) < 2e-5
