"""FBME-SBL numerical core on MI355X: UAMP + active-set EM plans, PCG path and
graph plans backed by hand-written HIP kernels (libfbme_hip.so).

Public names follow torch_sbl/__init__.py:1-22 of the reference for the FBME
hot path, including off-grid refinement (``refine_offgrid_fourier``, one HIP
launch per batch: 1-D and 2-D grids, at most 64 candidates and 64 refined
components).  The reference's autofocus and correlated-MMV research
extensions are not implemented here.
"""

import os as _os
import sys as _sys

_PKG_ROOT = _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__)))
if _PKG_ROOT not in _sys.path:
    _sys.path.insert(0, _PKG_ROOT)

from .version import PAPER_RELEASE_ID, __version__  # noqa: E402
from .config import SBLConfig, SBLResult  # noqa: E402
from .solver import clear_operator_cache, fast_sbl, sbl_1d, sbl_2d  # noqa: E402
from .operators import PrefixFourierOperator  # noqa: E402
from .cuda_graph import (  # noqa: E402
    CUDAGraphSBLPlan,
    RealtimeSBLPlan,
    make_cuda_graph_sbl_plan,
    make_realtime_sbl_plan,
)
from .hip_kernels import reset_triton_runtime_error  # noqa: E402
from .offgrid import (  # noqa: E402
    OffGridResult,
    continuous_fourier_dictionary,
    refine_offgrid_fourier,
)

__all__ = [
    "PAPER_RELEASE_ID", "__version__", "SBLConfig", "SBLResult",
    "clear_operator_cache", "fast_sbl", "sbl_1d", "sbl_2d",
    "PrefixFourierOperator", "OffGridResult", "continuous_fourier_dictionary",
    "refine_offgrid_fourier", "CUDAGraphSBLPlan", "RealtimeSBLPlan",
    "make_cuda_graph_sbl_plan", "make_realtime_sbl_plan",
    "reset_triton_runtime_error",
]
