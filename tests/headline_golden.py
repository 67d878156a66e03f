"""Headline-size reference fixtures (tests/golden/headline_*.npz): loading,
scene regeneration and the comparison both the oracle test and the GPU test
apply.

The fixtures were written by the reference's CPU Torch backend
(tests/golden/make_golden.py, ``headline``).  They hold the generator
arguments and the SHA-256 of ``y`` instead of ``y`` itself; ``load`` rebuilds
the scene and refuses a digest mismatch, so a change in the generator cannot
silently change what is compared.

Bar (the hot path's parity bar): relative L2 <= 1e-4 on x, identical
polished support and support mask, exchange count and iteration count;
noise, dof and the evidence trace to the float32 rounding of the reference.
The Stable gamma fixed-point residual is a maximum over atoms of relative
changes, so per-atom float32 differences pass into it unaveraged: it is held
to 1e-4 absolute or 1e-3 relative (the reference's own 0 dB run at 512x512
differs from the float64 oracle by 3.5e-4 relative there, with per-atom
posterior variances that differ by up to 2.3e-4).
"""

from __future__ import annotations

import hashlib
import os

import numpy as np

from fbme_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))

CASES_256 = ["headline_256x256_160x160_s1000", "headline_256x256_160x160_s1001",
             "headline_256x256_160x160_s1002"]
CASES_512 = ["headline_512x512_384x256_0db", "headline_512x512_384x256_10db",
             "headline_512x512_384x256_20db"]
CASES = [(c, p) for c in CASES_256 for p in ("fast", "stable")] + \
        [(c, "stable") for c in CASES_512]


def load(name):
    g = np.load(os.path.join(HERE, "golden", name + ".npz"))
    grid = tuple(int(v) for v in g["grid"])
    samples = tuple(int(v) for v in g["samples"])
    fn = scenes.ellipse_scene if str(g["generator"]) == "ellipse" else scenes.clustered_scene
    scene = fn(int(g["seed"]), grid, samples, k=int(g["k"]), snr_db=float(g["snr_db"]))
    digest = hashlib.sha256(np.ascontiguousarray(scene.y, np.complex64).tobytes()).hexdigest()
    if digest != str(g["y_sha256"]):
        raise AssertionError(f"{name}: regenerated scene differs from the fixture's input")
    return g, scene


def dense_x(g, profile, n):
    x = np.zeros(n, np.complex128)
    x[g[profile + "_x_index"]] = g[profile + "_x_value"]
    return x


def compare(g, profile, out, *, trace_rtol=1e-5, trace_atol=1e-2):
    """``out``: dict with x [N], posterior_variance [N], support_mask [N]
    (bool), noise_variance, effective_dof, evidence_trace, iterations and,
    for Stable, support_exchanges_accepted / fixed_budget_* fields."""
    p = profile + "_"
    n = int(np.prod(g["grid"]))
    x = np.asarray(out["x"]).reshape(-1)
    xr = dense_x(g, profile, n)
    rel = float(np.linalg.norm(x - xr) / np.linalg.norm(xr))
    assert rel < 1e-4, f"relative L2 {rel:.3e}"
    assert np.array_equal(np.flatnonzero(x), np.sort(g[p + "x_index"]))
    mask = np.asarray(out["support_mask"]).reshape(-1)
    assert np.array_equal(np.flatnonzero(mask), np.sort(g[p + "support_mask_index"]))
    var = np.asarray(out["posterior_variance"]).reshape(-1)
    idx = g[p + "x_index"]
    ref_var = g[p + "variance_on_support"]
    rel_var = float(np.linalg.norm(var[idx] - ref_var) / np.linalg.norm(ref_var))
    assert rel_var < 1e-4, f"posterior variance on the support: relative L2 {rel_var:.3e}"
    assert abs(float(np.sum(var, dtype=np.float64)) / float(g[p + "variance_sum"]) - 1) < 1e-4
    assert abs(float(out["noise_variance"]) / float(g[p + "noise_variance"]) - 1) < 1e-4
    assert abs(float(out["effective_dof"]) - float(g[p + "effective_dof"])) < 1e-2
    np.testing.assert_allclose(np.asarray(out["evidence_trace"]).reshape(-1),
                               g[p + "evidence_trace"].reshape(-1),
                               rtol=trace_rtol, atol=trace_atol)
    assert int(out["iterations"]) == int(g[p + "iterations"])
    if profile == "stable":
        assert int(out["support_exchanges_accepted"]) == int(g[p + "support_exchanges_accepted"])
        got = float(out["fixed_budget_gamma_residual"])
        want = float(g[p + "fixed_budget_gamma_residual"])
        assert abs(got - want) < max(1e-4, 1e-3 * abs(want))
        assert bool(out["fixed_budget_monotone"]) == bool(g[p + "fixed_budget_monotone"])
    return rel


# Generic-preset fixtures (tests/golden/make_golden.py, ``generic``):
# SBLConfig.fast_hybrid_uamp (make_cuda_graph_sbl_plan's default, with the
# truncation diagnostic) and SBLConfig.fast_uamp.
GENERIC_HYBRID = ["generic_hybrid_64x64_40x40", "generic_hybrid_96x80_60x48",
                  "generic_hybrid_256x256_160x160"]
GENERIC_UAMP = ["generic_uamp_64x64_40x40", "generic_uamp_96x80_60x48",
                "generic_uamp_256x256_160x160"]


def load_generic(name):
    g = np.load(os.path.join(HERE, "golden", name + ".npz"))
    grid = tuple(int(v) for v in g["grid"])
    samples = tuple(int(v) for v in g["samples"])
    fn = dict(uniform=scenes.uniform_scene, ellipse=scenes.ellipse_scene)[str(g["generator"])]
    scene = fn(int(g["seed"]), grid, samples, k=int(g["k"]), snr_db=float(g["snr_db"]))
    digest = hashlib.sha256(np.ascontiguousarray(scene.y, np.complex64).tobytes()).hexdigest()
    if digest != str(g["y_sha256"]):
        raise AssertionError(f"{name}: regenerated scene differs from the fixture's input")
    overrides = {k[len("override_"):]: g[k].item() for k in g.files if k.startswith("override_")}
    return g, scene, overrides


def compare_generic(g, out):
    """x to 1e-4 relative L2 with the identical polished support; posterior
    variance and precision on the support and summed over the grid; noise,
    dof, evidence; the truncation certificate (a maximum over the embedding
    spectrum, float32 in the reference) to 1e-4 relative."""
    n = int(np.prod(g["grid"]))
    x = np.asarray(out["x"]).reshape(-1)
    xr = np.zeros(n, np.complex128)
    xr[g["x_index"]] = g["x_value"]
    rel = float(np.linalg.norm(x - xr) / np.linalg.norm(xr))
    assert rel < 1e-4, f"relative L2 {rel:.3e}"
    assert np.array_equal(np.flatnonzero(x), np.sort(g["x_index"]))
    idx = g["x_index"]
    for field in ("variance", "precision"):
        v = np.asarray(out["posterior_variance" if field == "variance" else "precision"]).reshape(-1)
        ref = g[field + "_on_support"]
        assert float(np.linalg.norm(v[idx] - ref) / np.linalg.norm(ref)) < 1e-4, field
        assert abs(float(np.sum(v, dtype=np.float64)) / float(g[field + "_sum"]) - 1) < 1e-4, field
    assert abs(float(out["noise_variance"]) / float(g["noise_variance"]) - 1) < 1e-4
    assert abs(float(out["effective_dof"]) - float(g["effective_dof"])) < 1e-2
    np.testing.assert_allclose(np.asarray(out["evidence_trace"]).reshape(-1),
                               g["evidence_trace"].reshape(-1), rtol=2e-5, atol=0.1)
    assert int(out["iterations"]) == int(g["iterations"])
    if "truncation_delta_bound" in g.files:
        got = float(out["truncation_delta_bound"])
        want = float(g["truncation_delta_bound"])
        assert abs(got / want - 1) < 1e-4, (got, want)
        if "hard_pruning_accepted" in out:
            assert bool(out["hard_pruning_accepted"]) == bool(g["hard_pruning_accepted"])
    return rel
