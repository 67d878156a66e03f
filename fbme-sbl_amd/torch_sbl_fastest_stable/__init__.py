"""Locked Fast/Stable profiles of the FBME-SBL v1.1 core on MI355X
(torch_sbl_fastest_stable/__init__.py:1-19)."""

from .profiles import (
    CORE_RELEASE_ID,
    DUAL_RELEASE_ID,
    FastestSBLPlan,
    StableSBLPlan,
    make_fastest_plan,
    make_stable_plan,
)

__all__ = [
    "CORE_RELEASE_ID",
    "DUAL_RELEASE_ID",
    "FastestSBLPlan",
    "StableSBLPlan",
    "make_fastest_plan",
    "make_stable_plan",
]
