"""Host logic of the learned warm start (no GPU): the flat parameter layout
against the library's count and the oracle's initialisation, the committed
weights' metadata, and the stall criterion that sends a solve to the plain
fallback."""

import os
import types

import numpy as np
import torch

from oracle import cnn_oracle as cn


def test_parameter_layout_matches_library_and_oracle():
    from fbme_amd import _lib
    from torch_sbl import learned
    lib = _lib.load()
    p = learned.init_params(0)
    assert p.size == int(lib.fbme_cnn_param_count())
    layers = learned.unflatten(p)
    ref = cn.init_params(0)
    for (w, b), (rw, rb) in zip(layers, ref):
        np.testing.assert_allclose(w, rw, rtol=1e-6, atol=0)
        np.testing.assert_array_equal(b, rb)
    assert np.concatenate([np.concatenate([w.ravel(), b]) for w, b in layers]).tobytes() == p.tobytes()


def test_committed_weights_metadata():
    from torch_sbl import learned
    assert os.path.exists(learned.WEIGHTS)
    with np.load(learned.WEIGHTS) as z:
        assert z["params"].dtype == np.float32 and z["params"].size == learned.init_params().size
        assert np.all(np.isfinite(z["params"]))
        assert int(z["warm_iter"]) < int(z["target_iter"])
        assert tuple(z["grid"]) == (256, 256) and tuple(z["samples"]) == (160, 160)
        assert int(z["trajectories"]) >= 9000


def _solver(sums, noise):
    return types.SimpleNamespace(sums=torch.tensor([[s, 0.0, 0.0] for s in sums], dtype=torch.float64),
                                 noise=torch.tensor(noise, dtype=torch.float32))


def test_stall_criterion():
    from torch_sbl import learned
    cfg = types.SimpleNamespace(cg_max_iter=150)
    before = np.array([1.0, 2.0])
    assert not learned._stalled(_solver([1.5, 2.5], [0.1, 0.2]), 40, cfg, before)
    assert learned._stalled(_solver([1.5, 2.5], [0.1, 0.2]), 150, cfg, before)      # PCG budget exhausted
    assert learned._stalled(_solver([1.5, 2.5e3], [0.1, 0.2]), 40, cfg, before)    # variances blew up
    assert learned._stalled(_solver([1.5, 2.5], [0.1, float("nan")]), 40, cfg, before)
    assert learned._stalled(_solver([float("inf"), 2.5], [0.1, 0.2]), 40, cfg, before)
