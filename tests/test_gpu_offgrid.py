"""Off-grid refinement (refine_offgrid_fourier, torch_sbl/offgrid.py:70-226)
on the HIP kernel against the float64 oracle and the reference's golden runs,
on 2-D and 1-D grids.

The kernel evaluates the same variable-projection objective in FP64 through
closed-form atom inner products, so it matches the float64 oracle to
rounding (1e-7 bins, 1e-6 relative amplitude).  Model orders above 32
components keep the refinement matrices in a device workspace instead of
LDS; a scene refined either way gives the same bits.  Against the reference's own
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
          "offgrid_1d_128_96_b2_k4",
          # model orders above 32 (the refinement matrices in the workspace)
          "offgrid_128x128_80x80_k40", "offgrid_128x128_80x80_c128_k36", "offgrid_96x96_64x64_b2_c128_k48",
          "offgrid_1d_1024_640_c128_k40"]


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


def test_workspace_path_equals_lds_path(device):
    """fbme_offgrid_refine_ws with a workspace and fbme_offgrid_refine_dims
    (LDS only) on scenes with K <= 32: bitwise equal outputs."""
    import ctypes

    from fbme_amd import _lib, scenes
    lib = _lib.load()
    made = [scenes.offgrid_scene(1810 + i, (64, 64), (40, 40), k=7, snr_db=20.0) for i in range(3)]
    y = torch.from_numpy(np.stack([m[0].y for m in made])).to(device)
    x = torch.from_numpy(np.stack([m[0].x for m in made]).astype(np.complex64)).to(device)
    batch, rows = 3, 7
    eps, tiny = float(np.finfo(np.float32).eps), float(np.finfo(np.float32).tiny)

    def run(ws):
        f64 = dict(dtype=torch.float64, device=device)
        out = [torch.full((batch, rows, 2), -1.0, **f64), torch.full((batch, rows, 2), -1.0, **f64),
               torch.zeros((batch, rows), dtype=torch.complex128, device=device),
               torch.full((batch, rows, 2), -1.0, **f64), torch.zeros((batch, 2), **f64),
               torch.zeros((batch, 3), dtype=torch.int32, device=device)]
        stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
        args = [stream, batch, 64, 64, 40, 40, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()), 0,
                rows, rows, 0.1, 1.5, 15, 1e-4, 0.35, 1e-3, 1e-7, eps, tiny,
                *[ctypes.c_void_p(t.data_ptr()) for t in out], 2]
        if ws is None:
            _lib.check(lib.fbme_offgrid_refine_dims(*args), "fbme_offgrid_refine_dims")
        else:
            _lib.check(lib.fbme_offgrid_refine_ws(*args, ctypes.c_void_p(ws.data_ptr()), ws.numel()),
                       "fbme_offgrid_refine_ws")
        torch.cuda.synchronize()
        return out

    nbytes = int(lib.fbme_offgrid_workspace_bytes(batch))
    assert nbytes == batch * 397312
    ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
    a, b = run(None), run(ws)
    assert int(a[5][:, 0].min()) == rows and int(a[5][:, 1].min()) > 0
    for u, v in zip(a, b):
        assert torch.equal(u, v)
    # a workspace smaller than batch x 388 KiB is refused
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    rc = lib.fbme_offgrid_refine_ws(stream, batch, 64, 64, 40, 40, ctypes.c_void_p(y.data_ptr()),
                                    ctypes.c_void_p(x.data_ptr()), 0, 40, 64, 0.1, 1.5, 15, 1e-4, 0.35, 1e-3,
                                    1e-7, eps, tiny, *[ctypes.c_void_p(t.data_ptr()) for t in a], 2,
                                    ctypes.c_void_p(ws.data_ptr()), nbytes - 256)
    assert rc == _lib.ERR_ARG


def test_validation(device):
    y = torch.zeros((2, 40, 40), dtype=torch.complex64, device=device)
    x = torch.zeros((2, 64, 64), dtype=torch.complex64, device=device)
    with pytest.raises(ValueError):
        _run(y, x, (64, 64))                       # batch needs component_count
    with pytest.raises(ValueError):
        _run(y[0], x[0], (64, 64), relative_threshold=1.5)
    with pytest.raises(NotImplementedError):
        _run(y[0], x[0], (64, 64), max_components=100)
