"""CPU oracle for the FBME hot path -- TEST INFRASTRUCTURE ONLY.

This package restates the reference algorithm in float64 NumPy so that the
HIP implementation can be checked against it.  It is imported only by
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py``; the product path (``fbme-sbl_amd/``) never imports it and fails
loudly when its HIP library is missing.

Modules
-------
fbme_oracle   prefix DFT operator, FFT-UAMP sweep, active-set EM polish,
              fixed-budget diagnostics and result tail of the Fast/Stable plans
              (torch_sbl/solver.py:925-2393, 2726-2870, 3408-3496, 3591-3779,
              3961-4511; torch_sbl/triton_kernels.py:36-417).
cg_oracle     covariance-free PCG + Hutchinson path of ``sbl_2d`` (mode="cg")
              (torch_sbl/solver.py:1194-1221, 2991-3036, 5582-5848, 5979-6118;
              torch_sbl/operators.py:299-416).
dense_em      exact dense EM-SBL, the float64 reference solver
              (literature_reproduction/baselines.py:254-278 with
              torch_sbl/solver.py:2962-2990, operators.py:282-297, 418-442).
cnn_oracle    float64 reference of the learned preconditioner network's
              forward and backward passes (the project's own network; it has
              no counterpart in the reference).

Parity pins are documented in each module header.
"""
