"""Train the preconditioner network of the learned solver (GPU).

    python -m torch_sbl.train_precond --out precond.npz [--scenes 256]

The network emits the per-cell log-scale o of the grid-diagonal background
of the PCG preconditioner, P = B + F F^H with B^-1 = A diag(w) A^H,
w = exp(clamp(o)) / (N (noise + N gamma_T)) and F the 32 strongest atoms
(fbme_gdiag_weights); o = 0 is the closed form.

1. States: the plain general solver (16 Hutchinson probes, PCG tolerance
   1e-4) runs --em-iters EM iterations on synthetic ellipse scenes (256x256
   grid, 160x160 prefix, K=300, 10 dB) with training seeds
   fbme_amd.scenes.TRAIN_SEED_BASE + SEED_OFFSET + i (disjoint from the
   evaluation seeds and from the warm start's training scenes); every EM
   iteration's variances and noise, the state a PCG solve is preconditioned
   from, are kept.  With the warm-start network, the warm-started solver's
   states are kept too.
2. Loss: k PCG iterations are unrolled on the right-hand sides [y, 4
   Rademacher probes] of a state, preconditioned by P(o), and the loss is the
   mean over right-hand sides of log(||r_k||^2 / ||r_0||^2), the residual
   reduction the preconditioner buys in k steps.  Its gradient with respect
   to o is taken by torch's autograd over torch.fft; the network's forward
   and backward passes and Adam are the fbme_cnn_* kernels of libfbme_hip.so
   (fbme_cnn_output_grad feeds the loss gradient into fbme_cnn_backward).
   A frobenius-type loss ||P^-1 Q s - s||^2 was tried and gave worse
   preconditioners than the closed form: the unrolled residual is what
   tracks PCG iterations.
3. Evaluation on held-out scenes (evaluation seeds), one scene per solve:
   PCG iterations, scenes/s and NMSE of the plain solver (low-rank
   preconditioner, the reference's), the closed-form grid-diagonal
   background and the learned one, at --eval-iters EM iterations, and of the
   warm-started solver with the learned preconditioner.
The weights are written to --out (npz) and the metrics to --out + .json.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from fbme_amd import scenes  # noqa: E402
from torch_sbl import SBLConfig  # noqa: E402
from torch_sbl import learned, pcg  # noqa: E402

SEED_OFFSET = 50_000
RANK = 32


def solver_config(device, max_iter: int) -> SBLConfig:
    return SBLConfig(mode="cg", device=str(device), max_iter=max_iter, min_iter=1, tol=0.0,
                     hutchinson_probes=16, auto_probe_scaling=False, cg_tolerance=1e-4)


class States:
    """Solver states (gamma, noise) with their scene's data and backprojection."""

    def __init__(self):
        self.y, self.bp, self.bpe = [], [], []
        self.gamma, self.noise, self.scene = [], [], []

    def add_scenes(self, y, bp, bpe):
        base = sum(v.shape[0] for v in self.y)
        self.y.append(y)
        self.bp.append(bp)
        self.bpe.append(bpe)
        return base

    def finish(self):
        self.y = torch.cat(self.y)
        self.bp = torch.cat(self.bp)
        self.bpe = torch.cat(self.bpe)
        self.gamma = torch.cat(self.gamma)
        self.noise = torch.cat(self.noise)
        self.scene = torch.cat(self.scene)

    def __len__(self):
        return int(self.gamma.shape[0])


def collect(states: States, device, count: int, batch: int, grid, samples, k, snr_db, em_iters, warm_net, log):
    done = 0
    t0 = time.perf_counter()
    while done < count:
        nb = min(batch, count - done)
        sc = [scenes.ellipse_scene(scenes.TRAIN_SEED_BASE + SEED_OFFSET + done + i, grid, samples, k=k,
                                   snr_db=snr_db) for i in range(nb)]
        y = torch.from_numpy(np.stack([s.y for s in sc]).astype(np.complex64)).to(device)
        base = None
        for use_warm in ([False, True] if warm_net is not None else [False]):
            kept = []

            def record(s):
                ok = torch.isfinite(s.gamma).all(1) & torch.isfinite(s.noise) & (s.noise > 0)
                kept.append((s.gamma.clone(), s.noise.clone(), ok))

            def warm(s):
                nonlocal base
                if base is None:
                    base = states.add_scenes(s.y.clone(), s.bp.clone(), s.bpe.clone())
                if use_warm:
                    x = warm_net.features(s.bp, s.gamma, s.bpe, s.sums[:, 0].contiguous(), grid)
                    out = warm_net.forward(x)
                    learned._lib.check(warm_net.lib.fbme_cnn_apply_warm_start(
                        s.stream, s.b, s.n, learned._ptr(s.gamma), learned._ptr(out), learned._ptr(s.floor),
                        learned.CLAMP), "warm_start")
                    s.grid_sums(with_mean=False)
                record(s)

            def hook(s, iters, cfg, sum_before):
                record(s)
                return False

            iters = 3 if use_warm else em_iters
            pcg.sbl_cg(y, grid, solver_config(device, iters), batch_shape=(nb,), warm_start=warm, stall_check=hook)
            # the state after the last update preconditions no solve of this run; keep it anyway
            for g, nz, ok in kept:
                sel = torch.nonzero(ok).reshape(-1)
                states.gamma.append(g[sel])
                states.noise.append(nz[sel])
                states.scene.append(sel + base)
        done += nb
        log(f"[precond] states from {done}/{count} scenes ({time.perf_counter() - t0:.1f} s)")
    states.finish()


def torch_ops(grid, samples):
    n1, n2 = grid
    m1, m2 = samples
    n = n1 * n2

    def fwd(x):      # [..., N] -> [..., M]
        lead = x.shape[:-1]
        return torch.fft.fft2(x.reshape(*lead, n1, n2))[..., :m1, :m2].reshape(*lead, m1 * m2)

    def adj(v):      # [..., M] -> [..., N], N ifft2 of the zero-padded samples
        lead = v.shape[:-1]
        f = torch.zeros((*lead, n1, n2), dtype=v.dtype, device=v.device)
        f[..., :m1, :m2] = v.reshape(*lead, m1, m2)
        return torch.fft.ifft2(f).reshape(*lead, n) * n

    return fwd, adj


def columns(idx, val, grid, samples):
    """[S, 32, M] factor columns exp(-2 pi i (m1 k1 / N1 + m2 k2 / N2)) sqrt(gamma_k)."""
    n1, n2 = grid
    m1, m2 = samples
    k1 = idx // n2
    k2 = idx % n2
    a1 = torch.arange(m1, device=idx.device)
    a2 = torch.arange(m2, device=idx.device)
    ph1 = ((a1[None, None, :] * k1[..., None]) % n1).double() / n1      # [S, 32, M1]
    ph2 = ((a2[None, None, :] * k2[..., None]) % n2).double() / n2      # [S, 32, M2]
    u = torch.polar(torch.ones_like(ph1), -2 * math.pi * ph1)
    v = torch.polar(torch.ones_like(ph2), -2 * math.pi * ph2) * val.double().clamp_min(0).sqrt()[..., None]
    return (u[..., :, None] * v[..., None, :]).reshape(*idx.shape, m1 * m2).to(torch.complex64)


def unrolled_loss(o, gamma, noise, rhs, grid, samples, ops, steps, clamp):
    """Mean log residual reduction after `steps` PCG iterations on rhs
    [S, R, M] preconditioned by the grid-diagonal background with log-scale o
    [S, N] (the algebra of fbme_gdiag_weights / fbme_gdiag_system /
    fbme_lowrank_apply_split, in torch)."""
    fwd, adj = ops
    n = gamma.shape[1]
    val, idx = torch.topk(gamma, RANK, dim=1)
    gt = gamma.scatter(1, idx, 0.0)
    w0 = 1.0 / (n * (noise[:, None] + n * gt))
    w = (w0 * torch.exp(o.clamp(-clamp, clamp)))[:, None, :]                     # [S, 1, N]
    f = columns(idx, val, grid, samples)                                        # [S, 32, M]

    def binv(v):
        return fwd(w * adj(v))

    u = binv(f)                                                                 # [S, 32, M]
    c = torch.eye(RANK, dtype=u.dtype, device=u.device) + f.conj() @ u.transpose(1, 2)
    ci = torch.linalg.inv(0.5 * (c + c.conj().transpose(1, 2)))

    def pre(v):                                                                 # v [S, R, M]
        p = v @ u.conj().transpose(1, 2)                                        # U^H v  [S, R, 32]
        return binv(v) - (p @ ci.transpose(1, 2)) @ u

    def q(v):
        return noise[:, None, None] * v + fwd(gamma[:, None, :] * adj(v))

    r = rhs
    z = pre(r)
    d = z
    rho = (r.conj() * z).sum(-1).real
    r0 = (r.conj() * r).sum(-1).real
    for _ in range(steps):
        qd = q(d)
        a = rho / (d.conj() * qd).sum(-1).real
        r = r - a[..., None] * qd
        z = pre(r)
        nr = (r.conj() * z).sum(-1).real
        d = z + (nr / rho)[..., None] * d
        rho = nr
    rk = (r.conj() * r).sum(-1).real
    return torch.log(rk / r0).mean()


def train(net, st: States, grid, samples, steps, per_step, unroll, lr, probes, log):
    ops = torch_ops(grid, samples)
    dev = net.device
    n = st.gamma.shape[1]
    m = st.y.shape[1]
    gen = torch.Generator(device="cpu").manual_seed(0)
    losses, base_losses = [], []
    t0 = time.perf_counter()
    for step in range(steps):
        pick = torch.randperm(len(st), generator=gen)[:per_step].to(dev)
        sc = st.scene[pick]
        gamma = st.gamma[pick].contiguous()
        noise = st.noise[pick].contiguous()
        y = st.y[sc].reshape(per_step, m)
        s = (2 * torch.randint(0, 2, (per_step, probes, m), generator=gen) - 1).to(dev, torch.complex64)
        rhs = torch.cat([y[:, None, :], s], 1)
        gsum = gamma.double().sum(1)
        x = net.features(st.bp[sc].contiguous(), gamma, st.bpe[sc].contiguous(), gsum, noise, grid)
        out = net.forward(x).reshape(per_step, n)
        o = out.detach().clone().requires_grad_(True)
        loss = unrolled_loss(o, gamma, noise, rhs, grid, samples, ops, unroll, net.clamp)
        loss.backward()
        g = o.grad.reshape(per_step, *grid).float().contiguous()
        net.backward_from(x, g)
        cosine = 0.5 * (1 + math.cos(math.pi * step / steps))
        net.adam(lr * (0.05 + 0.95 * cosine))
        losses.append(float(loss.detach()))
        if step % 50 == 0 or step == steps - 1:
            with torch.no_grad():
                base = unrolled_loss(torch.zeros_like(o), gamma, noise, rhs, grid, samples, ops, unroll, net.clamp)
            base_losses.append(float(base))
            log(f"[precond] step {step}/{steps} loss {np.mean(losses[-50:]):.4f} "
                f"(closed form on this batch {float(base):.4f}) ({time.perf_counter() - t0:.1f} s)")
    return losses


def evaluate(device, net, warm_net, seeds, grid, samples, k, snr_db, em_iters, log):
    rows = {}

    def run(name, iters, precond=None, pnet=None, wnet=None):
        its, nm, times, fb = [], [], [], []
        cfg = solver_config(device, iters)
        for seed in seeds:
            sc = scenes.ellipse_scene(seed, grid, samples, k=k, snr_db=snr_db)
            y = torch.from_numpy(sc.y[None].astype(np.complex64)).to(device)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if wnet is None and pnet is None:
                res = pcg.sbl_cg(y, grid, dataclass_replace(cfg, return_history=True), batch_shape=(1,),
                                 precond=precond)
            else:
                res = learned.learned_sbl_2d(y, grid, wnet, dataclass_replace(cfg, return_history=True),
                                             precond_net=pnet)
                fb.append(res.history.get("learned_fallback", [0.0])[0])
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            its.append(float(np.sum(res.history["cg_iterations"])))
            nm.append(scenes.nmse_db(res.x.reshape(-1).cpu().numpy(), sc.x.reshape(-1)))
        rows[name] = dict(em_iterations=iters, pcg_iterations=float(np.mean(its)), nmse_db=float(np.mean(nm)),
                          scenes_per_s=float((len(seeds) - 1) / np.sum(times[1:])),
                          fallbacks=float(np.sum(fb)) if fb else 0.0)
        log(f"[precond] eval {name}: {rows[name]}")

    for iters in em_iters:
        run(f"plain_low_rank_{iters}", iters)
        run(f"grid_diagonal_closed_form_{iters}", iters, precond="grid_diagonal")
        run(f"grid_diagonal_learned_{iters}", iters, pnet=net)
    if warm_net is not None:
        run("warm_start_low_rank_3", 3, wnet=warm_net)
        run("warm_start_learned_precond_3", 3, wnet=warm_net, pnet=net)
    return rows


def dataclass_replace(cfg, **kw):
    import dataclasses
    return dataclasses.replace(cfg, **kw)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=learned.PRECOND_WEIGHTS)
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--em-iters", type=int, default=10)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--per-step", type=int, default=4)
    ap.add_argument("--unroll", type=int, default=10)
    ap.add_argument("--probes", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-scenes", type=int, default=16)
    ap.add_argument("--eval-iters", type=int, nargs="*", default=[3, 10])
    ap.add_argument("--no-warm", action="store_true", help="leave out the warm-started states")
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--samples", type=int, default=160)
    ap.add_argument("--k", type=int, default=300)
    ap.add_argument("--snr-db", type=float, default=10.0)
    ap.add_argument("--eval-only", default=None, help="evaluate these weights, no training")
    args = ap.parse_args(argv)
    device = torch.device("cuda", 0)
    grid, samples = (args.grid, args.grid), (args.samples, args.samples)

    def log(msg):
        print(msg, flush=True)

    warm_net = None if args.no_warm or not os.path.exists(learned.WEIGHTS) else learned.WarmStartNet.load(device)
    seeds = [scenes.EVAL_SEED_BASE + 400 + i for i in range(args.eval_scenes)]
    if args.eval_only:
        net = learned.PrecondNet.load(device, args.eval_only)
        rows = evaluate(device, net, warm_net, seeds, grid, samples, args.k, args.snr_db, args.eval_iters, log)
        log(json.dumps(dict(eval_seeds=[seeds[0], seeds[-1]], eval=rows)))
        return
    st = States()
    t0 = time.perf_counter()
    collect(st, device, args.scenes, args.batch, grid, samples, args.k, args.snr_db, args.em_iters, warm_net, log)
    t_collect = time.perf_counter() - t0
    log(f"[precond] {len(st)} states")
    net = learned.PrecondNet(device, seed=1)
    t0 = time.perf_counter()
    losses = train(net, st, grid, samples, args.steps, args.per_step, args.unroll, args.lr, args.probes, log)
    t_train = time.perf_counter() - t0
    net.save(args.out)
    rows = evaluate(device, net, warm_net, seeds, grid, samples, args.k, args.snr_db, args.eval_iters, log)
    meta = dict(scenes=args.scenes, states=len(st), steps=args.steps, per_step=args.per_step, unroll=args.unroll,
                probes=args.probes, lr=args.lr, collect_s=t_collect, train_s=t_train,
                loss_first=float(np.mean(losses[:50])), loss_last=float(np.mean(losses[-50:])),
                eval_seeds=[seeds[0], seeds[-1]], eval=rows)
    with open(args.out + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(json.dumps(meta))


if __name__ == "__main__":
    main()
