"""fresco.optimize (fresco/optimize/__init__.py:1-3 of the reference): the
Result dict, LBFGS_CPP and ModifiedFireCPP, run on the GPU."""
from fresco_amd.optimize import GradientOptimizer, LBFGS_CPP, ModifiedFireCPP, Result  # noqa: F401

__all__ = ["Result", "LBFGS_CPP", "ModifiedFireCPP", "GradientOptimizer"]
