"""ctypes binding of libfbme_hip.so (the C ABI declared in include/fbme_hip.h).

The library is loaded from this directory only.  When it is missing or fails
to load, every entry point raises: there is no CPU or PyTorch fallback.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_double, c_float, c_int,
                    c_longlong, c_size_t, c_uint, c_void_p)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "libfbme_hip.so")

ABI_VERSION = 9
# return codes (include/fbme_hip.h)
ERR_ARG = 1
ERR_HIP = 2

Q_WORKSPACE_BYTES = 0
Q_TOTAL_GRID = 1
Q_TOTAL_SAMPLES = 2
Q_MAX_ITERS = 3
Q_NSCENES = 4
Q_EVALUATIONS = 5
Q_LAUNCHES = 6
Q_R16_COL_ITEMS = 7
Q_R16_ROW_ITEMS = 8

STAGE_COL_ADJOINT = 0
STAGE_ROW = 1
STAGE_COL_FORWARD = 2
STAGE_OUTPUT = 3


class FbmePlanOptions(Structure):
    _fields_ = [
        ("uamp_damping", c_float), ("damping", c_float),
        ("min_variance", c_float), ("max_precision", c_float),
        ("blend", c_float), ("noise_fraction", c_float),
        ("margin_factor", c_float), ("jitter_multiplier", c_float),
        ("false_alarm", c_float), ("polish_size", c_int),
        ("polish_iterations", c_int), ("exchange_size", c_int),
        ("diagnostics", c_int), ("evidence_noise", c_int),
        ("row_lines", c_int), ("col_lines", c_int),
        ("iteration_offset", c_float), ("iteration_scale", c_float),
        ("min_iters", c_int), ("max_iters", c_int),
        ("hyper_shape", c_float), ("hyper_rate", c_float),
        ("noise_shape", c_float), ("noise_rate", c_float),
        ("polish_damping", c_float),
        ("fft_kernel", c_int),
        ("truncation_diagnostic", c_int), ("prune_tail", c_int),
        ("reinitialize", c_int), ("evidence_gate", c_int),
        ("fixed_noise", c_float),
    ]


class FbmeIO(Structure):
    _fields_ = [
        ("y", c_void_p), ("x", c_void_p), ("posterior_variance", c_void_p),
        ("precision", c_void_p), ("support_mask", c_void_p),
        ("noise", c_void_p), ("dof", c_void_p), ("nle", c_void_p),
        ("trace", c_void_p), ("exchanges", c_void_p), ("diag", c_void_p),
        ("status", c_void_p), ("store_uamp", c_int), ("truncation", c_void_p),
    ]


_LIB = None
_LOAD_ERROR: Exception | None = None

# name -> (restype, argtypes)
_SIGNATURES = {
    "fbme_abi_version": (c_int, []),
    "fbme_stream_copy": (c_int, [c_void_p, c_void_p, c_void_p, c_longlong, c_int]),
    "fbme_stream_fill": (c_int, [c_void_p, c_void_p, c_longlong, c_float, c_int]),
    "fbme_default_options": (None, [POINTER(FbmePlanOptions), c_int]),
    "fbme_plan_create": (c_void_p, [c_int, POINTER(c_int),
                                    POINTER(FbmePlanOptions), c_char_p, c_int]),
    "fbme_plan_create_masked": (c_void_p, [c_int, POINTER(c_int), POINTER(c_int),
                                           POINTER(c_int), POINTER(FbmePlanOptions),
                                           c_char_p, c_int]),
    "fbme_plan_set_mask": (c_int, [c_void_p, c_void_p, c_void_p]),
    "fbme_plan_destroy": (None, [c_void_p]),
    "fbme_plan_query": (c_longlong, [c_void_p, c_int]),
    "fbme_plan_scene_info": (c_int, [c_void_p, c_int, POINTER(c_longlong)]),
    "fbme_plan_bind": (c_int, [c_void_p, c_void_p, c_void_p]),
    "fbme_plan_run": (c_int, [c_void_p, POINTER(FbmeIO), c_float, c_void_p]),
    "fbme_plan_debug_pointer": (c_int, [c_void_p, c_int, POINTER(c_void_p)]),
    "fbme_plan_begin": (c_int, [c_void_p, c_void_p, c_void_p]),
    "fbme_plan_iterate": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "fbme_plan_finish": (c_int, [c_void_p, c_void_p, c_void_p]),
    "fbme_plan_run_stage": (c_int, [c_void_p, c_int, POINTER(FbmeIO), c_void_p]),
    "fbme_prefix_forward": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                    c_int, c_void_p, c_void_p, c_void_p]),
    "fbme_prefix_adjoint": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                    c_int, c_void_p, c_void_p, c_void_p]),
    "fbme_uamp_message_update": (c_int, [c_void_p, c_int, c_int, c_void_p,
                                         c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p,
                                         c_float]),
    "fbme_uamp_denoise": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_float, c_float, c_float,
                                  c_float, c_float, c_float, c_void_p]),
    "fbme_uamp_prior_update": (c_int, [c_void_p, c_int, c_int, c_void_p,
                                       c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_float, c_float, c_float,
                                       c_float]),
    "fbme_complex_residual_energy": (c_int, [c_void_p, c_int, c_int,
                                             c_void_p, c_void_p, c_void_p,
                                             c_void_p]),
    "fbme_uamp_noise_update": (c_int, [c_void_p, c_int, c_int, c_void_p,
                                       c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_float, c_float,
                                       c_float, c_int, c_float, c_void_p]),
    # general-solver (mode="cg") kernels
    "fbme_radix_code": (c_uint, [c_int]),
    "fbme_column_fft": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int,
                                c_void_p, c_void_p, c_int]),
    "fbme_dirichlet_table": (c_int, [c_int, c_int, c_void_p]),
    "fbme_twiddle_table": (c_int, [c_void_p, c_int, c_void_p]),
    "fbme_gamma_spectrum": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                    c_int, c_int, c_int, c_uint, c_uint,
                                    c_uint, c_uint, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "fbme_toeplitz_matvec": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                     c_int, c_int, c_uint, c_uint, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p]),
    "fbme_toeplitz_matvec_tiled": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                           c_int, c_int, c_uint, c_uint, c_void_p,
                                           c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_void_p, c_int, c_int]),
    "fbme_toeplitz_tiles": (c_int, [c_int, c_int, c_void_p, c_void_p]),
    "fbme_grid_matvec": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p]),
    "fbme_grid_matvec_dot": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p]),
    "fbme_lowrank_apply_dot": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                       c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p]),
    "fbme_lowrank_map": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "fbme_lowrank_apply_split": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                         c_void_p, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_gdiag_weights": (c_int, [c_void_p, c_int, c_longlong, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_float, c_void_p]),
    "fbme_lowrank_columns": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "fbme_gdiag_apply": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    "fbme_gdiag_system": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p]),
    "fbme_pcg_update_parts": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_int, c_void_p, c_void_p]),
    "fbme_pcg_direction_parts": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_int,
                                         c_void_p, c_void_p]),
    "fbme_pcg_update_parts_rr": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "fbme_lowrank_apply_fft_direction": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                                 c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                                 c_void_p]),
    "fbme_pcg_commit_valid": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "fbme_lowrank_apply_fft_dot": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                           c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_topk32_scratch_bytes": (c_size_t, [c_int, c_longlong]),
    "fbme_circulant_spectrum": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_uint, c_uint, c_void_p,
                                        c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_circulant_apply": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_uint, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p]),
    "fbme_topk_scratch_bytes": (c_size_t, [c_int, c_longlong, c_int]),
    "fbme_topk": (c_int, [c_void_p, c_int, c_longlong, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_reduced_posterior_large_bytes": (c_size_t, [c_int, c_int]),
    "fbme_reduced_posterior_large": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                             c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_void_p,
                                             c_size_t, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_topk32_noalloc": (c_int, [c_void_p, c_int, c_longlong, c_void_p, c_void_p, c_void_p]),
    "fbme_topk32": (c_int, [c_void_p, c_int, c_longlong, c_void_p, c_void_p,
                            c_void_p, c_void_p]),
    "fbme_lowrank_setup": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                   c_int, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p]),
    "fbme_lowrank_apply": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                   c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p]),
    "fbme_lane_dot": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p,
                              c_void_p, c_void_p]),
    "fbme_pcg_update": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p]),
    "fbme_pcg_direction": (c_int, [c_void_p, c_int, c_int, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_int]),
    "fbme_sub": (c_int, [c_void_p, c_longlong, c_void_p, c_void_p, c_void_p]),
    "fbme_pcg_check": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_float, c_int]),
    "fbme_masked_copy": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p,
                                 c_void_p]),
    "fbme_hutchinson": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p]),
    "fbme_cg_em": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                           c_void_p, c_void_p, c_void_p, c_void_p, c_float,
                           c_float, c_float, c_float, c_int, c_void_p,
                           c_void_p, c_void_p]),
    "fbme_cg_init": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p,
                             c_void_p, c_void_p, c_float, c_float, c_float,
                             c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_relevance_sum": (c_int, [c_void_p, c_int, c_int, c_void_p,
                                   c_void_p, c_void_p, c_void_p]),
    "fbme_reciprocal": (c_int, [c_void_p, c_longlong, c_void_p, c_void_p]),
    "fbme_grid_sums": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p]),
    # learned warm start (convolutional network on BF16 MFMA)
    "fbme_cnn_param_count": (c_longlong, []),
    "fbme_cnn_weight_count": (c_longlong, []),
    "fbme_cnn_prepare": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_cnn_features": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p]),
    "fbme_cnn_features_precond": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_cnn_output_grad": (c_int, [c_void_p, c_longlong, c_void_p, c_void_p]),
    "fbme_cnn_forward": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_cnn_grad_rows": (c_longlong, [c_int]),
    "fbme_cnn_part_floats": (c_longlong, [c_int, c_int, c_int]),
    "fbme_cnn_loss": (c_int, [c_void_p, c_longlong, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p]),
    "fbme_cnn_backward": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p]),
    "fbme_cnn_adam": (c_int, [c_void_p, c_longlong, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_float, c_int]),
    "fbme_cnn_apply_warm_start": (c_int, [c_void_p, c_int, c_int, c_void_p,
                                          c_void_p, c_void_p, c_float]),
    "fbme_reduced_posterior": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                       c_int, c_int, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p,
                                       c_float, c_void_p, c_void_p, c_void_p,
                                       c_void_p]),
    # exact dense posterior (mode="dense")
    "fbme_dense_workspace_bytes": (c_size_t, [c_int, c_int, c_int, c_int, c_int, c_int]),
    "fbme_dense_posterior": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                     c_int, c_int, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_double, c_double,
                                     c_void_p, c_size_t, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p]),
    # off-grid refinement
    "fbme_offgrid_refine": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                    c_int, c_void_p, c_void_p, c_int, c_int,
                                    c_int, c_float, c_float, c_int, c_double,
                                    c_double, c_double, c_double, c_double,
                                    c_double, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "fbme_offgrid_refine_dims": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                         c_int, c_void_p, c_void_p, c_int, c_int,
                                         c_int, c_float, c_float, c_int, c_double,
                                         c_double, c_double, c_double, c_double,
                                         c_double, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_int]),
    "fbme_offgrid_workspace_bytes": (c_size_t, [c_int]),
    "fbme_offgrid_refine_ws": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                       c_int, c_void_p, c_void_p, c_int, c_int,
                                       c_int, c_float, c_float, c_int, c_double,
                                       c_double, c_double, c_double, c_double,
                                       c_double, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_int,
                                       c_void_p, c_size_t]),
    # joint sparsity of the general solver
    "fbme_joint_init": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_float, c_int, c_void_p,
                                c_void_p]),
    "fbme_hutchinson_shared": (c_int, [c_void_p, c_int, c_int, c_int,
                                       c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p]),
    "fbme_cg_em_joint": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_float, c_float, c_float, c_float,
                                 c_int, c_void_p, c_void_p, c_void_p]),
    "fbme_joint_noise_update": (c_int, [c_void_p, c_int, c_int, c_void_p,
                                        c_int, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p, c_float,
                                        c_float, c_float, c_int, c_float,
                                        c_void_p]),
    # observation masks of the general solver
    "fbme_mask_spectrum": (c_int, [c_void_p, c_int, c_int, c_int, c_int,
                                   c_void_p, c_void_p, c_void_p]),
    "fbme_mask_lanes_dot": (c_int, [c_void_p, c_int, c_int, c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_lowrank_setup_masked": (c_int, [c_void_p, c_int, c_int, c_int,
                                          c_int, c_int, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p]),
    "fbme_reduced_posterior_masked": (c_int, [c_void_p, c_int, c_int, c_int,
                                              c_int, c_int, c_void_p,
                                              c_void_p, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_float,
                                              c_void_p, c_void_p, c_void_p,
                                              c_void_p]),
    "fbme_uamp_noise_update_masked": (c_int, [c_void_p, c_int, c_int,
                                              c_void_p, c_int, c_void_p,
                                              c_void_p, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_float,
                                              c_float, c_float, c_int,
                                              c_float, c_void_p]),
    "fbme_fixed_noise": (c_int, [c_void_p, c_int, c_float, c_void_p, c_void_p, c_int]),
    "fbme_hutchinson_used": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_spatial_init": (c_int, [c_void_p, c_int, c_int, c_int, c_float, c_int, c_float, c_void_p, c_void_p,
                                  c_void_p]),
    "fbme_spatial_update": (c_int, [c_void_p, c_int, c_int, c_int, c_float, c_int, c_float, c_float, c_float,
                                    c_float, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    # Nystrom preconditioner and the stored-column apply
    "fbme_stored_apply": (c_int, [c_void_p, c_int, c_int, c_longlong, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_nystrom_part_bytes": (c_size_t, [c_int]),
    "fbme_nystrom_omega": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                   c_void_p]),
    "fbme_nystrom_setup": (c_int, [c_void_p, c_int, c_longlong, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fbme_nystrom_vectors": (c_int, [c_void_p, c_int, c_longlong, c_int, c_void_p, c_void_p, c_void_p]),
}

EXPORTED = tuple(_SIGNATURES)


def load():
    """Load the HIP library (raising with the reason when it cannot)."""
    global _LIB, _LOAD_ERROR
    if _LIB is not None:
        return _LIB
    if _LOAD_ERROR is not None:
        raise RuntimeError(
            f"libfbme_hip.so failed to load: {_LOAD_ERROR}") from _LOAD_ERROR
    if not os.path.exists(LIB_PATH):
        _LOAD_ERROR = FileNotFoundError(LIB_PATH)
        raise RuntimeError(
            f"{LIB_PATH} is missing; build it with "
            "`python fbme-sbl_amd/fbme_amd/build.py` (hipcc, gfx950)")
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as exc:
        _LOAD_ERROR = exc
        raise RuntimeError(f"libfbme_hip.so failed to load: {exc}") from exc
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.fbme_abi_version() != ABI_VERSION:
        raise RuntimeError("libfbme_hip.so ABI version mismatch")
    _LIB = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


def default_options(stable: bool) -> FbmePlanOptions:
    opt = FbmePlanOptions()
    load().fbme_default_options(ctypes.byref(opt), 1 if stable else 0)
    return opt
