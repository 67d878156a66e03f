"""MI355X runtime of the FBME hot path: native plans over libfbme_hip.so,
batched uneven workloads, autotuning, synthetic scenes and the learned
preconditioner."""
