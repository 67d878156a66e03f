"""Algorithmic HBM bytes of the FBME plan, per scene and per kernel.

Two views are kept:

* ``stagewise_scene_bytes`` -- the stage-wise model of the whole solve used to
  compare against the reference: every stage reads its inputs once and
  writes its outputs once, FFT intermediates are free.  Per UAMP iteration
  64 M + 56 N bytes; Fast = (56 + 64 T_u) M + (93 + 56 T_u) N, Stable adds
  8 M + 12 N (M = M1 M2 samples, N = N1 N2 grid points, complex64 = 8 B,
  float32 = 4 B).
* ``kernel_bytes`` -- the unique bytes one launch of each fused pass must
  move in this implementation's two-pass FFT design, including the
  [N1, M2] intermediate that necessarily round-trips between the column and
  row passes.  This is the ``achieved`` numerator of the roofline.
"""

from __future__ import annotations


def stagewise_scene_bytes(n: int, m: int, t_u: int, stable: bool = False) -> int:
    fast = (56 + 64 * t_u) * m + (93 + 56 * t_u) * n
    return fast + (8 * m + 12 * n if stable else 0)


def kernel_bytes(n1: int, n2: int, m1: int, m2: int) -> dict[str, int]:
    n = n1 * n2
    m = m1 * m2
    mid = n1 * m2
    return {
        # residual y - A mu and old message in; new message and [N1, M2] out
        "col_adjoint": 16 * m + 8 * m + 8 * mid,
        # [N1, M2] in; gamma, mean, variance in and out; [N1, M2] out
        "row": 8 * mid + 16 * n + 16 * n + 8 * mid,
        # [N1, M2] in, data in, residual out
        "col_forward": 8 * mid + 8 * m + 8 * m,
        # x, posterior variance, precision, support mask out
        "output": 8 * n + 4 * n + 4 * n + n,
    }
