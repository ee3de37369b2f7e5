"""CombinedPotential and PythonPotential.

CombinedPotential mirrors fresco/potentials/combine_potentials.pyx:10-30 over
source/fresco/combine_potentials.hpp:16-59; as_cpp_potential mirrors
fresco/potentials/_pythonpotential.pyx:58-65 (any object with getEnergyGradient
/ get_energy_gradient becomes a potential the device optimizers can drive).
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from ._base import BasePotential


class CombinedPotential(BasePotential):
    """Sum of potentials over the same coordinates."""

    def __init__(self, potentials=()):
        self._pots = []
        for p in potentials:
            self._pots.append(as_cpp_potential(p))
        if not self._pots:
            raise ValueError("CombinedPotential needs at least one potential")
        self._rebuild()

    def add_potential(self, potential):
        self._pots.append(as_cpp_potential(potential))
        self._rebuild()

    def _rebuild(self):
        lib = _lib.load()
        old = self._h
        arr = (ctypes.c_void_p * len(self._pots))(*[p._h for p in self._pots])
        h = ctypes.c_void_p()
        _lib.check(lib.fresco_combined_create(arr, len(self._pots), ctypes.byref(h)))
        if old:
            lib.fresco_pot_destroy(old)
        self._init_handle(h)


class PythonPotential(BasePotential):
    """Wraps a host-side Python potential (getEnergyGradient / get_energy_gradient
    returning (energy, grad)) as a device potential; x is copied to the host
    for each call."""

    def __init__(self, pot, ndof):
        self.pot = pot
        fn = getattr(pot, "getEnergyGradient", None) or getattr(pot, "get_energy_gradient")
        efn = getattr(pot, "getEnergy", None) or getattr(pot, "get_energy", None)
        n = int(ndof)

        def cb(xp, gp, ep, user):
            try:
                x = np.ctypeslib.as_array(xp, shape=(n,)).copy()
                if gp:
                    e, g = fn(x)
                    np.ctypeslib.as_array(gp, shape=(n,))[:] = np.asarray(g, dtype=np.float64).ravel()
                else:
                    e = efn(x) if efn is not None else fn(x)[0]
                ep[0] = float(e)
                return 0
            except Exception:  # reported as a RuntimeError by the library
                import traceback

                traceback.print_exc()
                return 1

        self._cb = _lib.EG_CALLBACK(cb)
        h = ctypes.c_void_p()
        _lib.check(_lib.load().fresco_callback_create(n, ctypes.cast(self._cb, ctypes.c_void_p), None, ctypes.byref(h)))
        self._init_handle(h)


def as_cpp_potential(pot, verbose=False, ndof=None):
    """Return a device potential: the object itself if it already is one,
    otherwise a PythonPotential around it (ndof inferred from `ndof`)."""
    if isinstance(pot, BasePotential):
        return pot
    if ndof is None:
        raise ValueError("a Python potential needs ndof (pass it through the optimizer's x0)")
    if verbose:
        print("making a python potential from", pot)
    return PythonPotential(pot, ndof)
