"""Public entry point of the locked minimum-latency profile."""

from ..profiles import FastestSBLPlan, make_fastest_plan

__all__ = ["FastestSBLPlan", "make_fastest_plan"]
