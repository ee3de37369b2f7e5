"""Device optimizers (mirror of fresco.optimize)."""
from ._opt import GradientOptimizer, LBFGS_CPP, ModifiedFireCPP
from .result import Result

__all__ = ["GradientOptimizer", "LBFGS_CPP", "ModifiedFireCPP", "Result"]
