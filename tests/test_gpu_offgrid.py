"""Off-grid refinement (refine_offgrid_fourier, torch_sbl/offgrid.py:70-226)
on the HIP kernel against the float64 oracle and the reference's golden runs,
on 2-D and 1-D grids.

The kernel evaluates the same variable-projection objective in FP64 through
closed-form atom inner products, so it matches the float64 oracle to
rounding (1e-7 bins, 1e-6 relative amplitude).  Against the reference's own
runs, which compute in float32 for complex64 data, the bar is that of the
oracle: identical candidate peaks and model order, positions within 5e-5
bins, amplitudes within 5e-5 relative, residual norms within 1e-5, position
standard deviations within 1e-3 relative; the complex128 run also has equal
iteration counts and convergence flags.
"""

import numpy as np
import pytest
import torch

from conftest import golden_path
from oracle import offgrid_oracle as oo

pytestmark = pytest.mark.gpu

GOLDEN = ["offgrid_64x64_40x40_c128", "offgrid_64x64_40x40_c64", "offgrid_64x64_40x40_b2_k5",
          "offgrid_128x96_80x60", "offgrid_1d_256_160_c64", "offgrid_1d_512_200_c128",
          "offgrid_1d_128_96_b2_k4"]


def _run(y, x, grid, **kw):
    from torch_sbl import refine_offgrid_fourier
    return refine_offgrid_fourier(y, x, grid, **kw)


@pytest.mark.parametrize("name", GOLDEN)
def test_offgrid_matches_reference_golden_and_oracle(device, name):
    print("golden", name, flush=True)
    g = np.load(golden_path(name + ".npz"))
    grid = tuple(int(v) for v in g["grid"])
    count = int(g["component_count"])
    count = None if count < 0 else count
    exact = g["y"].dtype == np.complex128
    res = _run(torch.from_numpy(g["y"]).to(device), torch.from_numpy(g["x"]).to(device), grid,
               component_count=count)
    ftype = np.float64 if exact else np.float32
    orc = oo.refine_offgrid(g["y"], g["x"], grid, component_count=count,
                            eps=float(np.finfo(ftype).eps), tiny=float(np.finfo(ftype).tiny))
    d = len(grid)          # 1-D grids: one position coordinate per component
    pos = res.positions.cpu().numpy().reshape(len(orc), -1, d)
    amp = res.amplitudes.cpu().numpy().reshape(len(orc), -1)
    std = res.position_std.cpu().numpy().reshape(len(orc), -1, d)
    init = res.initial_positions.cpu().numpy().reshape(len(orc), -1, d)
    norm = res.residual_norm.cpu().numpy().reshape(-1)
    for i, o in enumerate(orc):
        np.testing.assert_array_equal(init[i], g["initial_positions"][i])
        # FP64 kernel against the FP64 oracle
        np.testing.assert_allclose(pos[i], o["positions"], atol=1e-7)
        np.testing.assert_allclose(amp[i], o["amplitudes"], rtol=1e-6, atol=1e-6)
        # against the reference run
        np.testing.assert_allclose(pos[i], g["positions"][i], atol=1e-9 if exact else 5e-5)
        np.testing.assert_allclose(amp[i], g["amplitudes"][i], rtol=5e-5, atol=5e-5)
        np.testing.assert_allclose(std[i], g["position_std"][i], rtol=1e-3)
        assert norm[i] == pytest.approx(float(g["residual_norm"][i]), rel=1e-5)
        if exact:
            assert int(res.iterations.reshape(-1)[i]) == int(g["iterations"][i])
            assert bool(res.converged.reshape(-1)[i]) == bool(g["converged"][i])


def test_offgrid_256_grid_vs_oracle(device):
    """256 x 256 grid, 160 x 160 prefix, 20 off-grid sources at 20 dB,
    model order by BIC."""
    from fbme_amd import scenes
    sc, true_pos, _ = scenes.offgrid_scene(1700, (256, 256), (160, 160), k=20, snr_db=20.0)
    res = _run(torch.from_numpy(sc.y).to(device), torch.from_numpy(sc.x.astype(np.complex64)).to(device),
               (256, 256))
    orc = oo.refine_offgrid(sc.y, sc.x.astype(np.complex64), (256, 256))[0]
    pos = res.positions.cpu().numpy()
    print("256 offgrid: K", pos.shape[0], "iterations", int(res.iterations), flush=True)
    assert pos.shape == orc["positions"].shape == (20, 2)
    np.testing.assert_allclose(pos, orc["positions"], atol=1e-6)
    err = np.abs(np.sort(pos, 0) - np.sort(true_pos, 0)).max()
    assert err < 0.05


def test_batch_equals_single_items(device):
    from fbme_amd import scenes
    made = [scenes.offgrid_scene(1800 + i, (64, 64), (40, 40), k=5, snr_db=20.0) for i in range(4)]
    y = torch.from_numpy(np.stack([m[0].y for m in made])).to(device)
    x = torch.from_numpy(np.stack([m[0].x for m in made]).astype(np.complex64)).to(device)
    batch = _run(y, x, (64, 64), component_count=5)
    for i in range(4):
        single = _run(y[i], x[i], (64, 64), component_count=5)
        assert torch.equal(batch.positions[i], single.positions)
        assert torch.equal(batch.amplitudes[i], single.amplitudes)


def test_fast_sbl_offgrid_refinement(device):
    """fast_sbl(..., offgrid_refinement=True) refines the plan's estimate
    (solver.py:2139-2153): sources 4+ bins apart recovered within 0.05 bins."""
    from fbme_amd import scenes
    from torch_sbl import SBLConfig, fast_sbl
    sc, true_pos, _ = scenes.offgrid_scene(1900, (64, 64), (40, 40), k=6, snr_db=25.0)
    cfg = SBLConfig.paper_hybrid_uamp(device=str(device), offgrid_refinement=True)
    res = fast_sbl(torch.from_numpy(sc.y).to(device), (64, 64), config=cfg)
    assert res.offgrid is not None
    pos = res.offgrid.positions.cpu().numpy()
    assert pos.shape[1] == 2
    for p in true_pos:
        d = np.abs(pos - p)
        d = np.minimum(d, 64 - d)
        assert np.min(np.hypot(d[:, 0], d[:, 1])) < 0.05


def test_validation(device):
    y = torch.zeros((2, 40, 40), dtype=torch.complex64, device=device)
    x = torch.zeros((2, 64, 64), dtype=torch.complex64, device=device)
    with pytest.raises(ValueError):
        _run(y, x, (64, 64))                       # batch needs component_count
    with pytest.raises(ValueError):
        _run(y[0], x[0], (64, 64), relative_threshold=1.5)
    with pytest.raises(NotImplementedError):
        _run(y[0], x[0], (64, 64), max_components=100)
