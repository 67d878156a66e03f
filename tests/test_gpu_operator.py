"""Pruned 2-D FFT kernels (fbme_prefix_forward / fbme_prefix_adjoint) against
the float64 oracle, and the operator identities at full size."""

import numpy as np
import pytest
import torch

from oracle import fbme_oracle as fo

pytestmark = pytest.mark.gpu

SHAPES = [((32, 32), (24, 24)), ((40, 32), (20, 16)), ((64, 64), (40, 40)),
          ((96, 80), (60, 48)), ((192, 192), (77, 130)), ((256, 256), (160, 160)),
          ((512, 512), (384, 256)), ((128, 64), (128, 64)), ((1, 240), (1, 100)),
          ((30, 45), (1, 1))]


def _rand(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


@pytest.mark.parametrize("grid,samples", SHAPES)
def test_prefix_forward_adjoint_vs_oracle(device, grid, samples):
    from torch_sbl import PrefixFourierOperator
    print("shape", grid, samples, flush=True)
    op = PrefixFourierOperator(grid, samples, device=device)
    rng = np.random.default_rng(11)
    b = 3
    x = _rand(rng, b, op.n_grid)
    v = _rand(rng, b, op.n_samples)
    xt = torch.from_numpy(x.astype(np.complex64)).to(device)
    vt = torch.from_numpy(v.astype(np.complex64)).to(device)
    fwd = op.forward_flat(xt).cpu().numpy()
    adj = op.adjoint_flat(vt).cpu().numpy()
    ref_f = fo.prefix_forward(x.astype(np.complex64), grid, samples)
    ref_a = fo.prefix_adjoint(v.astype(np.complex64), grid, samples)
    # float32 FFT: error ~ eps * log2(N) relative to the L2 norm
    assert np.linalg.norm(fwd - ref_f) / np.linalg.norm(ref_f) < 2e-6
    assert np.linalg.norm(adj - ref_a) / np.linalg.norm(ref_a) < 2e-6


def test_adjoint_identity_and_orthogonal_rows(device):
    from torch_sbl import PrefixFourierOperator
    grid, samples = (256, 256), (160, 160)
    op = PrefixFourierOperator(grid, samples, device=device)
    g = torch.Generator(device="cpu").manual_seed(5)
    x = torch.randn(2, op.n_grid, dtype=torch.complex64, generator=g).to(device)
    v = torch.randn(2, op.n_samples, dtype=torch.complex64, generator=g).to(device)
    lhs = torch.sum(torch.conj(v) * op.forward_flat(x)).cpu().numpy()
    rhs = torch.sum(torch.conj(op.adjoint_flat(v)) * x).cpu().numpy()
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    # A A^H = N I
    back = op.forward_flat(op.adjoint_flat(v)).cpu().numpy()
    ref = op.n_grid * v.cpu().numpy()
    assert np.linalg.norm(back - ref) / np.linalg.norm(ref) < 2e-6


def test_multi_rhs_layout(device):
    from torch_sbl import PrefixFourierOperator
    grid, samples = (40, 32), (20, 16)
    op = PrefixFourierOperator(grid, samples, device=device)
    rng = np.random.default_rng(3)
    d = _rand(rng, 2, op.n_samples, 5).astype(np.complex64)
    out = op.adjoint(torch.from_numpy(d).to(device)).cpu().numpy()
    for r in range(5):
        ref = fo.prefix_adjoint(d[:, :, r], grid, samples)
        assert np.linalg.norm(out[:, :, r] - ref) / np.linalg.norm(ref) < 2e-6
