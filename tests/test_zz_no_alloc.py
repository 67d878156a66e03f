"""fbme_topk32_noalloc (candidates in the library's fixed per-GPU scratch, no
allocation) against fbme_topk32 with a caller workspace, through the C ABI:
two calls in a row, the second with a smaller input, bit-exact indices and
values (integer selection work).  Inputs beyond the scratch's maximum are
rejected before any launch."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def test_topk32_noalloc_matches_workspace_entry_point(device):
    from fbme_amd import _lib
    lib = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    limit = 4 << 20
    for b, n in [(8, 262144), (3, 40000)]:
        print("noalloc top-32", b, n, flush=True)
        need = int(lib.fbme_topk32_scratch_bytes(b, n))
        assert 0 < need <= limit
        rng = np.random.default_rng(b * n)
        x = (rng.integers(0, 256, size=(b, n)) / 256.0).astype(np.float32)
        xt = torch.from_numpy(x).to(device)
        cand = torch.empty(need // 8, dtype=torch.int64, device=device)
        idx_ref = torch.empty((b, 32), dtype=torch.int32, device=device)
        val_ref = torch.empty((b, 32), dtype=torch.float32, device=device)
        _lib.check(lib.fbme_topk32(stream, b, n, _ptr(xt), _ptr(cand), _ptr(idx_ref), _ptr(val_ref)), "topk32")
        idx = torch.full((b, 32), -1, dtype=torch.int32, device=device)
        val = torch.full((b, 32), -1.0, dtype=torch.float32, device=device)
        _lib.check(lib.fbme_topk32_noalloc(stream, b, n, _ptr(xt), _ptr(idx), _ptr(val)), "topk32_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(idx, idx_ref)
        assert torch.equal(val, val_ref)
        want = np.argsort(-x, axis=1, kind="stable")[:, :32]
        np.testing.assert_array_equal(idx.cpu().numpy(), want)
    # beyond the scratch maximum: an argument error, nothing launched
    b, n = 64, 1 << 20
    assert int(lib.fbme_topk32_scratch_bytes(b, n)) > limit
    dummy = torch.empty(1, dtype=torch.float32, device=device)
    out_i = torch.empty(1, dtype=torch.int32, device=device)
    assert lib.fbme_topk32_noalloc(stream, b, n, _ptr(dummy), _ptr(out_i), _ptr(dummy)) == _lib.ERR_ARG
