"""Synthetic scene generators for the FBME workloads.

The reference ships no generator (README.md:58-65, RTX4090_EXAMPLE.md:28), so
the workloads are defined here from the observation model of README.md:58:

    data = fft2(x)[:M1, :M2] + noise,  noise ~ CN(0, sigma^2)
    SNR  = 10 log10(||A x||^2 / (M sigma^2))

All randomness comes from ``numpy.random.Generator(PCG64(seed))`` so a scene is
reproducible from its seed on any host.  Evaluation seeds start at 1000 and
training seeds at 1_000_000, so the two sets never overlap.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

EVAL_SEED_BASE = 1000
TRAIN_SEED_BASE = 1_000_000


@dataclass
class Scene:
    """One synthetic scene: ground truth, observations and metadata."""

    x: np.ndarray          # [N1, N2] complex128 ground truth
    y: np.ndarray          # [M1, M2] complex64 noisy prefix observations
    noise_variance: float  # sigma^2 of the complex AWGN
    grid: tuple[int, int]
    samples: tuple[int, int]
    snr_db: float
    seed: int

    @property
    def support(self) -> np.ndarray:
        return np.flatnonzero(self.x.reshape(-1) != 0)


def prefix_dft(x: np.ndarray, samples: tuple[int, int]) -> np.ndarray:
    """Unnormalised 2-D DFT restricted to its leading M1 x M2 block."""
    m1, m2 = samples
    return np.fft.fft2(x.astype(np.complex128))[:m1, :m2]


def _observe(rng, x, samples, snr_db, seed, grid):
    clean = prefix_dft(x, samples)
    m = clean.size
    sigma2 = float(np.sum(np.abs(clean) ** 2) / (m * 10.0 ** (snr_db / 10.0)))
    noise = np.sqrt(sigma2 / 2.0) * (
        rng.standard_normal(clean.shape) + 1j * rng.standard_normal(clean.shape)
    )
    y = (clean + noise).astype(np.complex64)
    return Scene(x=x, y=y, noise_variance=sigma2, grid=tuple(grid),
                 samples=tuple(samples), snr_db=float(snr_db), seed=int(seed))


def uniform_scene(seed: int, grid=(32, 32), samples=(24, 24), k=20,
                  snr_db=20.0) -> Scene:
    """K nonzeros at uniform positions, amplitudes CN(0, 1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = grid[0] * grid[1]
    x = np.zeros(n, dtype=np.complex128)
    pos = rng.choice(n, size=min(k, n), replace=False)
    x[pos] = np.sqrt(0.5) * (rng.standard_normal(pos.size)
                             + 1j * rng.standard_normal(pos.size))
    return _observe(rng, x.reshape(grid), samples, snr_db, seed, grid)


def _ellipse_cells(rng, grid, area_fraction, center=None, scale=1.0):
    n1, n2 = grid
    area = area_fraction * n1 * n2 * scale
    aspect = rng.uniform(1.0, 2.5)
    a = np.sqrt(area * aspect / np.pi)
    b = area / (np.pi * a)
    theta = rng.uniform(0.0, np.pi)
    if center is None:
        center = (rng.uniform(0.3, 0.7) * n1, rng.uniform(0.3, 0.7) * n2)
    r1, r2 = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
    d1 = r1 - center[0]
    d2 = r2 - center[1]
    u = d1 * np.cos(theta) + d2 * np.sin(theta)
    v = -d1 * np.sin(theta) + d2 * np.cos(theta)
    inside = (u / a) ** 2 + (v / b) ** 2 <= 1.0
    return np.flatnonzero(inside.reshape(-1))


def _lognormal_amplitudes(rng, count):
    mag = np.exp(rng.standard_normal(count))
    phase = rng.uniform(0.0, 2.0 * np.pi, count)
    return mag * np.exp(1j * phase)


def ellipse_scene(seed: int, grid=(256, 256), samples=(160, 160), k=300,
                  snr_db=10.0, area_fraction=0.2) -> Scene:
    """K point scatterers inside one random ellipse covering ~20% of the grid;
    log-normal (sigma=1) magnitudes with uniform phase."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cells = _ellipse_cells(rng, grid, area_fraction)
    pos = rng.choice(cells, size=min(k, cells.size), replace=False)
    x = np.zeros(grid[0] * grid[1], dtype=np.complex128)
    x[pos] = _lognormal_amplitudes(rng, pos.size)
    return _observe(rng, x.reshape(grid), samples, snr_db, seed, grid)


def clustered_scene(seed: int, grid=(512, 512), samples=(384, 256), k=1200,
                    snr_db=10.0, clusters=6, isolated_fraction=0.1) -> Scene:
    """Six clustered ellipses plus 10% isolated points (log-normal magnitudes)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = grid[0] * grid[1]
    k_iso = int(round(isolated_fraction * k))
    k_clu = k - k_iso
    chosen = set()
    per = [k_clu // clusters + (1 if i < k_clu % clusters else 0)
           for i in range(clusters)]
    for count in per:
        cells = _ellipse_cells(rng, grid, 0.02)
        cells = np.array([c for c in cells if c not in chosen], dtype=np.int64)
        take = rng.choice(cells, size=min(count, cells.size), replace=False)
        chosen.update(int(c) for c in take)
    free = np.setdiff1d(np.arange(n), np.fromiter(chosen, dtype=np.int64))
    iso = rng.choice(free, size=k_iso, replace=False)
    pos = np.concatenate([np.fromiter(chosen, dtype=np.int64), iso])
    x = np.zeros(n, dtype=np.complex128)
    x[pos] = _lognormal_amplitudes(rng, pos.size)
    return _observe(rng, x.reshape(grid), samples, snr_db, seed, grid)


def joint_scenes(seed: int, count=3, grid=(32, 32), samples=(24, 24), k=12,
                 snr_db=20.0) -> list[Scene]:
    """`count` snapshots sharing one support of K uniform positions, each with
    its own CN(0, 1) amplitudes and noise (the multiple-measurement-vector
    model of joint sparsity)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = grid[0] * grid[1]
    pos = rng.choice(n, size=min(k, n), replace=False)
    out = []
    for i in range(count):
        x = np.zeros(n, dtype=np.complex128)
        x[pos] = np.sqrt(0.5) * (rng.standard_normal(pos.size) + 1j * rng.standard_normal(pos.size))
        out.append(_observe(rng, x.reshape(grid), samples, snr_db, seed + i, grid))
    return out


def offgrid_scene(seed: int, grid=(64, 64), samples=(40, 40), k=6, snr_db=25.0,
                  min_separation=4.0) -> tuple[Scene, np.ndarray, np.ndarray]:
    """K point sources at continuous (off-grid) positions, at least
    `min_separation` bins apart (periodic), amplitudes CN(0, 1) rescaled to
    modulus in [0.5, 1.5].  Returns the scene (x holds each source's
    amplitude at its nearest cell, plus 1e-3-level clutter elsewhere, as a
    stand-in for an on-grid estimate), the positions [K, 2] and the
    amplitudes [K]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n1, n2 = grid
    pos = []
    while len(pos) < k:
        p = np.array([rng.uniform(0, n1), rng.uniform(0, n2)])
        ok = True
        for q in pos:
            d = np.abs(p - q)
            d = np.minimum(d, np.array(grid) - d)
            if np.hypot(*d) < min_separation:
                ok = False
        if ok:
            pos.append(p)
    pos = np.stack(pos)
    amp = rng.standard_normal(k) + 1j * rng.standard_normal(k)
    amp = amp / np.abs(amp) * rng.uniform(0.5, 1.5, k)
    m1, m2 = samples
    mm1, mm2 = np.meshgrid(np.arange(m1), np.arange(m2), indexing="ij")
    ph = mm1.reshape(-1, 1) * pos[None, :, 0] / n1 + mm2.reshape(-1, 1) * pos[None, :, 1] / n2
    clean = (np.exp(-2j * np.pi * ph) @ amp).reshape(samples)
    sigma2 = float(np.sum(np.abs(clean) ** 2) / (clean.size * 10.0 ** (snr_db / 10.0)))
    y = clean + np.sqrt(sigma2 / 2.0) * (rng.standard_normal(samples) + 1j * rng.standard_normal(samples))
    x = 1e-3 * (rng.standard_normal(grid) + 1j * rng.standard_normal(grid))
    cells = np.round(pos).astype(int) % np.array(grid)
    x[cells[:, 0], cells[:, 1]] = amp
    sc = Scene(x=x, y=y.astype(np.complex64), noise_variance=sigma2, grid=tuple(grid),
               samples=tuple(samples), snr_db=float(snr_db), seed=int(seed))
    return sc, pos, amp


def offgrid_scene_1d(seed: int, n=256, m=160, k=6, snr_db=25.0,
                     min_separation=4.0) -> tuple[Scene, np.ndarray, np.ndarray]:
    """offgrid_scene on a 1-D grid of n points sampled at its first m: x [n],
    y [m], positions [K, 1], amplitudes [K]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = []
    while len(pos) < k:
        p = rng.uniform(0, n)
        if all(min(abs(p - q), n - abs(p - q)) >= min_separation for q in pos):
            pos.append(p)
    pos = np.array(pos)
    amp = rng.standard_normal(k) + 1j * rng.standard_normal(k)
    amp = amp / np.abs(amp) * rng.uniform(0.5, 1.5, k)
    clean = np.exp(-2j * np.pi * np.outer(np.arange(m), pos) / n) @ amp
    sigma2 = float(np.sum(np.abs(clean) ** 2) / (m * 10.0 ** (snr_db / 10.0)))
    y = clean + np.sqrt(sigma2 / 2.0) * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
    x = 1e-3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    x[np.round(pos).astype(int) % n] = amp
    sc = Scene(x=x, y=y.astype(np.complex64), noise_variance=sigma2, grid=(n,), samples=(m,),
               snr_db=float(snr_db), seed=int(seed))
    return sc, pos[:, None], amp


def uneven_batch(seed: int, count=64, grids=(64, 128, 192, 256),
                 ratio=(0.3, 0.8), k_fraction=(0.005, 0.03),
                 snr_db=(5.0, 25.0)) -> list[Scene]:
    """Scenes with grids drawn from {64^2,128^2,192^2,256^2}, per-axis sampling
    ratio U[0.3, 0.8], K ~ U[0.5%, 3%] of the cells, SNR ~ U[5, 25] dB."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scenes = []
    for i in range(count):
        g = int(rng.choice(grids))
        m1 = max(1, int(round(rng.uniform(*ratio) * g)))
        m2 = max(1, int(round(rng.uniform(*ratio) * g)))
        k = max(1, int(round(rng.uniform(*k_fraction) * g * g)))
        snr = float(rng.uniform(*snr_db))
        sub_seed = int(rng.integers(0, 2**31 - 1))
        scenes.append(uniform_scene(sub_seed, grid=(g, g), samples=(m1, m2),
                                    k=k, snr_db=snr))
    return scenes


def nmse_db(estimate: np.ndarray, truth: np.ndarray) -> float:
    """10 log10(||x_hat - x||^2 / ||x||^2)."""
    err = np.sum(np.abs(estimate.astype(np.complex128) - truth) ** 2)
    return float(10.0 * np.log10(err / max(np.sum(np.abs(truth) ** 2), 1e-300)))


def support_f1(estimate_support, truth_support) -> float:
    est = set(int(i) for i in estimate_support)
    tru = set(int(i) for i in truth_support)
    if not est and not tru:
        return 1.0
    tp = len(est & tru)
    if tp == 0:
        return 0.0
    p = tp / len(est)
    r = tp / len(tru)
    return 2 * p * r / (p + r)
