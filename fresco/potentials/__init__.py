"""fresco.potentials (fresco/potentials/__init__.py:2-14 of the reference):
the four structure-factor potentials, the pair potentials and the
utility potentials, evaluated on the GPU."""
from fresco_amd.potentials import (  # noqa: F401
    BasePotential,
    CombinedPotential,
    InversePower,
    NUwNU,
    NUwU,
    NUwUBatch,
    PythonPotential,
    UwNU,
    UwU,
    WCAPotential,
    as_cpp_potential,
)

# _pythonpotential.pyx:31: the reference's name for a Python potential
# wrapped for its C++ optimizers
CppPotentialWrapper = PythonPotential

__all__ = ["UwU", "NUwU", "UwNU", "NUwNU", "InversePower", "WCAPotential", "BasePotential", "CombinedPotential",
           "PythonPotential", "CppPotentialWrapper", "as_cpp_potential", "NUwUBatch"]
