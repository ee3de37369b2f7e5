"""Structure-factor potentials on the GPU (mirror of fresco.potentials)."""
from ._base import BasePotential
from .nuwu import NUwU

__all__ = ["BasePotential", "NUwU"]
