"""Write float64-oracle fixtures for workloads the oracle needs minutes for.

    python tests/golden/make_oracle_fixtures.py

The general solver (mode='cg') at 512x512 / 384x256 with 32 probes takes the
float64 CG oracle (oracle/cg_oracle.py) about 200 s on 8 host cores, too long
to recompute inside a GPU test.  This script runs it once and stores its
outputs with the generator arguments and the SHA-256 of the input, so the
test regenerates the scene and compares against the stored result.
"""

from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "fbme-sbl_amd")]

from fbme_amd import scenes  # noqa: E402
from oracle import cg_oracle as co  # noqa: E402

CG_512 = dict(name="oracle_cg_512x512_384x256", seed=1000, grid=(512, 512), samples=(384, 256),
              k=1200, snr_db=10.0, probes=32, em_iterations=3)


def cg_512() -> None:
    c = CG_512
    sc = scenes.clustered_scene(c["seed"], c["grid"], c["samples"], k=c["k"], snr_db=c["snr_db"])
    cfg = co.CGConfig(max_iter=c["em_iterations"], hutchinson_probes=c["probes"], auto_probe_scaling=False)
    r = co.sbl_cg(sc.y[None], c["grid"], cfg)
    digest = hashlib.sha256(np.ascontiguousarray(sc.y, np.complex64).tobytes()).hexdigest()
    np.savez_compressed(
        os.path.join(HERE, c["name"] + ".npz"),
        seed=np.array(c["seed"]), grid=np.array(c["grid"]), samples=np.array(c["samples"]),
        k=np.array(c["k"]), snr_db=np.array(c["snr_db"]), probes=np.array(c["probes"]),
        em_iterations=np.array(c["em_iterations"]), y_sha256=np.array(digest),
        x=r["x"].reshape(-1).astype(np.complex64),
        posterior_variance=r["posterior_variance"].reshape(-1).astype(np.float32),
        noise_variance=np.asarray(r["noise_variance"], np.float64).reshape(-1),
        effective_dof=np.asarray(r["effective_dof"], np.float64).reshape(-1),
        iterations=np.array(r["iterations"]), cg_iterations=np.asarray(r["cg_iterations"]))
    print("wrote", c["name"], "pcg iterations", r["cg_iterations"],
          "nmse", scenes.nmse_db(r["x"].reshape(-1), sc.x.reshape(-1)))


if __name__ == "__main__":
    cg_512()
