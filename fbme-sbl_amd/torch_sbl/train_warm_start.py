"""Train the learned warm start on trajectories of the plain solver (GPU).

    python -m torch_sbl.train_warm_start --out weights.npz [--scenes 10000]

1. Trajectories: the plain general solver (mode='cg', 16 Hutchinson probes,
   PCG tolerance 1e-4) runs on synthetic ellipse scenes (256x256 grid,
   160x160 prefix, K=300, 10 dB by default) with training seeds
   fbme_amd.scenes.TRAIN_SEED_BASE + i, disjoint from the evaluation seeds.
   Each scene gives the network input (log|A^H y|^2, log gamma_0) and the
   target log(gamma_t / gamma_0), gamma_t being the plain solver's variances
   after --target-iter EM iterations.  Scenes whose plain solve is unstable
   (non-finite state or noise inflated more than tenfold) are dropped.
2. Training: Adam on the mean squared error with a cosine learning rate,
   minibatches of scenes, on the BF16 MFMA kernels of libfbme_hip.so.
3. Evaluation on held-out scenes (evaluation seeds), one scene per solve:
   the plain solver at the reference budget (--plain-iter EM iterations), the
   plain solver at the warm-start budget (--warm-iter) and at the target
   iteration, and the warm-started solver at --warm-iter; EM and PCG
   iterations, scenes/s and NMSE for each.
The weights (with target_iter and warm_iter) are written to --out (npz) and
the metrics to --out + .json.
"""

from __future__ import annotations

import argparse
import json
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


def make_batch(seed0: int, count: int, grid, samples, k: int, snr_db: float):
    batch = [scenes.ellipse_scene(seed0 + i, grid, samples, k=k, snr_db=snr_db) for i in range(count)]
    return batch, np.stack([s.y for s in batch]).astype(np.complex64)


def solver_config(device, max_iter: int) -> SBLConfig:
    """The plain solver of the learned comparison: fixed EM budget."""
    return SBLConfig(mode="cg", device=str(device), max_iter=max_iter, min_iter=1, tol=0.0,
                     hutchinson_probes=16, auto_probe_scaling=False, cg_tolerance=1e-4)


def collect(net, device, scenes_total: int, batch: int, grid, samples, k, snr_db, target_iter, log):
    xs, ts = [], []
    done = 0
    t0 = time.perf_counter()
    while done < scenes_total:
        n = min(batch, scenes_total - done)
        sc, y = make_batch(scenes.TRAIN_SEED_BASE + done, n, grid, samples, k, snr_db)
        yt = torch.from_numpy(y).to(device)
        x, target, res, valid = learned.trajectory_pairs(yt, grid, solver_config(device, target_iter), net)
        noise0 = np.array([s.noise_variance for s in sc])
        noise = res.noise_variance.reshape(-1).cpu().numpy()
        keep = valid.cpu().numpy() & (noise < 10.0 * np.maximum(noise0, 1e-30))
        sel = torch.from_numpy(np.flatnonzero(keep)).to(device)
        xs.append(x[sel].clone())
        ts.append(target[sel].clone())
        del x, target, res
        done += n
        log(f"[train] trajectories {done}/{scenes_total} kept {int(keep.sum())}/{n} "
            f"({time.perf_counter() - t0:.1f} s)")
    return torch.cat(xs), torch.cat(ts)


def train(net, x, t, steps: int, minibatch: int, lr: float, log):
    n = x.shape[0]
    g = torch.Generator(device="cpu").manual_seed(0)
    losses = []
    t0 = time.perf_counter()
    for step in range(1, steps + 1):
        idx = torch.randint(0, n, (minibatch,), generator=g).to(x.device)
        cur = lr * 0.5 * (1.0 + np.cos(np.pi * step / steps))
        loss = net.train_step(x[idx].contiguous(), t[idx].contiguous(), cur)
        if step % 250 == 0 or step == 1:
            losses.append((step, float(loss.item())))
            log(f"[train] step {step} loss {losses[-1][1]:.4f} ({time.perf_counter() - t0:.1f} s)")
    return losses


def _solve(kind, net, y, grid, device, iters):
    cfg = solver_config(device, iters)
    if kind == "learned":
        return learned.learned_sbl_2d(y, grid, net, cfg)
    return pcg.sbl_cg(y, grid, cfg, batch_shape=(1,))


def evaluate(net, device, count: int, grid, samples, k, snr_db, runs, log):
    """runs: [(name, kind, em_iterations)]; one scene per solve."""
    sc, y = make_batch(scenes.EVAL_SEED_BASE, count, grid, samples, k, snr_db)
    truth = np.stack([s.x.reshape(-1) for s in sc])
    out = {}
    yt0 = torch.from_numpy(y[:1]).to(device)
    for _, kind, iters in runs:  # warm-up (library load, workspaces)
        _solve(kind, net, yt0, grid, device, min(iters, 2))
    for name, kind, iters in runs:
        em, cg_its, nmse, secs, fallback = [], [], [], [], []
        for i in range(count):
            yt = torch.from_numpy(y[i:i + 1]).to(device)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = _solve(kind, net, yt, grid, device, iters)
            torch.cuda.synchronize()
            secs.append(time.perf_counter() - t0)
            em.append(res.iterations)
            cg_its.append(float(np.sum(res.history["cg_iterations"])))
            nmse.append(scenes.nmse_db(res.x.reshape(-1).cpu().numpy(), truth[i]))
            fallback.append(res.history.get("learned_fallback", [0.0])[0])
        out[name] = dict(kind=kind, em_iterations=float(np.mean(em)), pcg_iterations=float(np.mean(cg_its)),
                         seconds_per_scene=float(np.mean(secs)), scenes_per_s=float(1.0 / np.mean(secs)),
                         nmse_db=float(np.mean(nmse)), median_nmse_db=float(np.median(nmse)),
                         fallbacks=float(np.sum(fallback)), per_scene_nmse_db=[float(v) for v in nmse])
        log(f"[eval] {name}: {json.dumps({k2: v for k2, v in out[name].items() if not k2.startswith('per_')})}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--scenes", type=int, default=10000)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--samples", type=int, default=160)
    ap.add_argument("--k", type=int, default=300)
    ap.add_argument("--snr", type=float, default=10.0)
    ap.add_argument("--target-iter", type=int, default=10)
    ap.add_argument("--warm-iter", type=int, default=3)
    ap.add_argument("--plain-iter", type=int, default=100)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--minibatch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval", type=int, default=16)
    ap.add_argument("--extra-iters", default="2,5", help="further budgets evaluated for both solvers")
    args = ap.parse_args(argv)
    device = torch.device("cuda", 0)
    grid, samples = (args.grid, args.grid), (args.samples, args.samples)

    def log(msg):
        print(msg, flush=True)

    net = learned.WarmStartNet(device, seed=0)
    t0 = time.perf_counter()
    x, t = collect(net, device, args.scenes, args.batch, grid, samples, args.k, args.snr, args.target_iter, log)
    t_collect = time.perf_counter() - t0
    log(f"[train] {x.shape[0]} training trajectories")
    t0 = time.perf_counter()
    losses = train(net, x, t, args.steps, args.minibatch, args.lr, log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    net.target_iter, net.warm_iter = args.target_iter, args.warm_iter
    net.save(args.out, grid=np.array(grid), samples=np.array(samples), k=args.k, snr_db=args.snr,
             trajectories=x.shape[0], steps=args.steps, target_iter=args.target_iter, warm_iter=args.warm_iter)
    log("[train] wrote " + args.out)
    del x, t
    metrics = dict(grid=list(grid), samples=list(samples), k=args.k, snr_db=args.snr,
                   target_iter=args.target_iter, warm_iter=args.warm_iter,
                   scenes=args.scenes, collect_seconds=t_collect, train_seconds=t_train,
                   steps=args.steps, minibatch=args.minibatch, lr=args.lr, losses=losses)
    runs = [("plain_reference_budget", "plain", args.plain_iter),
            ("plain_target_iter", "plain", args.target_iter),
            ("plain_warm_budget", "plain", args.warm_iter),
            ("learned_warm_budget", "learned", args.warm_iter)]
    for it in (int(v) for v in args.extra_iters.split(",") if v):
        runs += [(f"plain_{it}", "plain", it), (f"learned_{it}", "learned", it)]
    metrics["evaluation"] = evaluate(net, device, args.eval, grid, samples, args.k, args.snr, runs, log)
    with open(args.out + ".json", "w") as fh:
        json.dump(metrics, fh, indent=1)
    log("[train] wrote " + args.out + ".json")


if __name__ == "__main__":
    main()
