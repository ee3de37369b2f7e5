"""Pair potentials: InversePower and WCAPotential.

Python surface of fresco/potentials/inversepower_potential.pyx:14-83 and
wca_potential.pyx:13-60 (constructor arguments, pickling through them) over
the device implementation of source/fresco/inversepower_potential.hpp and
wca.hpp.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from ..distances import Distance
from ._base import BasePotential

_KIND_IP = 0
_KIND_WCA = 1


def default_ncellx_scale(radii, boxv):
    """The reference's default (distance_utils.py get_ncellsx_scale) only
    tunes its OpenMP load balance; the GPU cell lists size their cells from
    the interaction range, so the value is recorded and otherwise unused."""
    return 1.0


class _PairPotential(BasePotential):
    _kind = None

    def _create(self, p1, eps, radii, boxv, use_cell_lists, ncellx_scale, balance_omp, method):
        self.boxv = np.ascontiguousarray(np.asarray(boxv, dtype=np.float64).ravel())
        self.radii = np.ascontiguousarray(np.asarray(radii, dtype=np.float64).ravel())
        self.use_cell_lists = bool(use_cell_lists)
        self.ncellx_scale = default_ncellx_scale(self.radii, self.boxv) if ncellx_scale is None else ncellx_scale
        self.balance_omp = balance_omp
        self.method = method
        self.ndim = self.boxv.size
        self.natoms = self.radii.size
        self.eps = float(eps)
        m = Distance(method).value if not isinstance(method, Distance) else method.value
        if m == Distance.CARTESIAN.value and self.use_cell_lists:
            raise NotImplementedError("cell lists need periodic boundaries")
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(
            lib.fresco_pair_create(
                self._kind, float(p1), self.eps, _lib.dptr(self.radii), self.natoms, self.ndim,
                _lib.dptr(self.boxv), m, int(self.use_cell_lists), float(self.ncellx_scale), ctypes.byref(h),
            )
        )
        self._init_handle(h)


class InversePower(_PairPotential):
    """E = sum_{i<j} eps/a (1 - r_ij/(r_i + r_j))^a for r_ij < r_i + r_j
    (inversepower_potential.hpp:29-58), cartesian or periodic (minimum image),
    all pairs or cell lists."""

    _kind = _KIND_IP

    def __init__(self, a, eps, radii, boxv, use_cell_lists=False, ncellx_scale=None, balance_omp=True,
                 method=Distance.PERIODIC):
        self.a = float(a)
        self._create(a, eps, radii, boxv, use_cell_lists, ncellx_scale, balance_omp, method)

    def __reduce__(self):
        return (InversePower, (self.a, self.eps, self.radii, self.boxv, self.use_cell_lists, self.ncellx_scale,
                               self.balance_omp, self.method))


class WCAPotential(_PairPotential):
    """E = sum_{i<j} 4 eps ((sigma/r)^12 - (sigma/r)^6) + eps for
    r < 2^(1/6) sigma (wca.hpp:34-63); radii do not enter the energy."""

    _kind = _KIND_WCA

    def __init__(self, sigma, eps, radii, boxv, use_cell_lists=False, ncellx_scale=None, balance_omp=True,
                 method=Distance.PERIODIC):
        self.sigma = float(sigma)
        self._create(sigma, eps, radii, boxv, use_cell_lists, ncellx_scale, balance_omp, method)

    def __reduce__(self):
        return (WCAPotential, (self.sigma, self.eps, self.radii, self.boxv, self.use_cell_lists, self.ncellx_scale,
                               self.balance_omp, self.method))
