"""Structure-factor potentials on the GPU (mirror of fresco.potentials)."""
from ._base import BasePotential
from .batch import NUwUBatch
from .combine import CombinedPotential, PythonPotential, as_cpp_potential
from .nuwnu import NUwNU
from .nuwu import NUwU
from .pair import InversePower, WCAPotential
from .uwnu import UwNU
from .uwu import UwU

__all__ = ["BasePotential", "CombinedPotential", "PythonPotential", "as_cpp_potential", "NUwU", "NUwNU", "UwU", "UwNU", "InversePower", "WCAPotential", "NUwUBatch"]
