"""What the trained warm-start network gains over the plain general solver.

Held-out scenes (seeds disjoint from training and from the training run's own
evaluation) of the training distribution: 256x256 grid, 160x160 prefix, K=300
ellipse scatterers, 10 dB.  The plain solver (mode='cg', 16 probes, PCG
tolerance 1e-4) and the warm-started solver run the same small EM budget;
the network must buy NMSE at that budget, and reach the plain solver's
trajectory quality at the iteration it imitates with fewer PCG iterations.
The fallback to the plain solver on a stalled PCG is checked to return the
plain solver's result exactly.
"""

import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GRID, SAMPLES, K, SNR = (256, 256), (160, 160), 300, 10.0
SEEDS = range(1200, 1206)


def _scenes():
    from fbme_amd import scenes
    sc = [scenes.ellipse_scene(s, GRID, SAMPLES, k=K, snr_db=SNR) for s in SEEDS]
    return sc, np.stack([s.y for s in sc]).astype(np.complex64)


def _cfg(device, iters, **kw):
    from torch_sbl import SBLConfig
    return SBLConfig(mode="cg", device=str(device), max_iter=iters, min_iter=1, tol=0.0,
                     hutchinson_probes=16, auto_probe_scaling=False, cg_tolerance=1e-4, **kw)


@pytest.fixture(scope="module")
def net(device):
    from torch_sbl import learned
    assert os.path.exists(learned.WEIGHTS), "trained weights missing: " + learned.WEIGHTS
    return learned.WarmStartNet.load(device)


def _run(kind, net, device, y, iters):
    from torch_sbl import learned, pcg
    cfg = _cfg(device, iters)
    out = []
    for i in range(y.shape[0]):
        yt = torch.from_numpy(y[i:i + 1]).to(device)
        r = learned.learned_sbl_2d(yt, GRID, net, cfg) if kind == "learned" else pcg.sbl_cg(
            yt, GRID, cfg, batch_shape=(1,))
        out.append(r)
    return out


def test_warm_start_gains_over_plain(device, net):
    from fbme_amd import scenes
    sc, y = _scenes()
    truth = [s.x.reshape(-1) for s in sc]

    def stats(rs):
        nm = [scenes.nmse_db(r.x.reshape(-1).cpu().numpy(), t) for r, t in zip(rs, truth)]
        its = [float(np.sum(r.history["cg_iterations"])) for r in rs]
        return float(np.mean(nm)), float(np.mean(its))

    budget = net.warm_iter
    plain_b = stats(_run("plain", net, device, y, budget))
    plain_t = stats(_run("plain", net, device, y, net.target_iter))
    rs = _run("learned", net, device, y, budget)
    learned_b = stats(rs)
    print(f"budget {budget}: plain {plain_b}, learned {learned_b}; plain at {net.target_iter}: {plain_t}",
          flush=True)
    assert all(r.history["learned_fallback"] == [0.0] for r in rs)
    # NMSE in dB: lower is better.  Gain at the same EM budget:
    assert learned_b[0] < plain_b[0] - 0.5
    # the imitated trajectory point's quality with far fewer PCG iterations
    assert learned_b[0] < plain_t[0] + 0.15
    assert learned_b[1] < 0.5 * plain_t[1]


def test_fallback_returns_plain_solution(device, net):
    """A PCG budget of 2 iterations stalls every warm-started solve: the
    batch is solved again from the plain initialisation."""
    from torch_sbl import learned, pcg
    _, y = _scenes()
    yt = torch.from_numpy(y[:2]).to(device)
    cfg = _cfg(device, 2, cg_max_iter=2)
    r = learned.learned_sbl_2d(yt, GRID, net, cfg)
    p = pcg.sbl_cg(yt, GRID, cfg, batch_shape=(2,))
    assert r.history["learned_fallback"] == [1.0]
    assert torch.equal(r.x, p.x)
    assert torch.equal(r.noise_variance, p.noise_variance)
