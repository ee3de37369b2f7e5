"""Structure-factor potentials on the GPU (mirror of fresco.potentials)."""
from ._base import BasePotential
from .combine import CombinedPotential, PythonPotential, as_cpp_potential
from .nuwu import NUwU

__all__ = ["BasePotential", "CombinedPotential", "PythonPotential", "as_cpp_potential", "NUwU"]
