"""NUwUBatch: many independent NUwU patterns evaluated together.

Each pattern is an NUwU problem (fresco/potentials/nuwu.pyx:10-23: radii,
Sk, V on a (2K'+1)^d mode grid; shared box L); the batch evaluates all of
them with one sort, one spreading and one interpolation launch and batched
FFTs (source/fresco/nuwu.hpp:152-317 per pattern).  No reference counterpart:
the reference evaluates one pattern per object.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from ._base import BasePotential, precision_code


class NUwUBatch(BasePotential):
    """patterns: sequence of (radii, Sk, V) per pattern (1D or 2D, d = len(L)).

    x is the concatenation of the patterns' coordinate vectors (pattern p at
    ``offsets[p] : offsets[p + 1]``); the energy is the sum of the pattern
    energies, ``pattern_energies()`` returns them one by one.  ``precision``:
    "fp32" (FP32 grids, default) or "fp64" (FP64 grids), both with the
    one-transform adjoint gradient.
    """

    def __init__(self, patterns, L, eps=1e-6, *, precision="fp32", sigma=None):
        self.L = np.ascontiguousarray(np.asarray(L, dtype=np.float64).ravel())
        self.ndim = len(self.L)
        self.eps = float(eps)
        self.precision = precision
        self.sigma = sigma
        radii, Sk, V, npts, nmodes = [], [], [], [], []
        for r, s, v in patterns:
            r = np.asarray(r, dtype=np.float64).ravel()
            s = np.asarray(s, dtype=np.float64).ravel()
            v = np.asarray(v, dtype=np.float64).ravel()
            if len(s) != len(v):
                raise ValueError("Sk and V differ in size")
            radii.append(r)
            Sk.append(s)
            V.append(v)
            npts.append(len(r))
            nmodes.append(len(v))
        self.npatterns = len(npts)
        self.npts = np.asarray(npts, dtype=np.int64)
        self.nmodes_per_pattern = np.asarray(nmodes, dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.npts * self.ndim)])
        self._radii = np.ascontiguousarray(np.concatenate(radii))
        self._Sk = np.ascontiguousarray(np.concatenate(Sk))
        self._V = np.ascontiguousarray(np.concatenate(V))
        code = precision_code(precision)
        if code == _lib.PREC_FP64_PLAIN:
            code = _lib.PREC_FP64_ADJOINT
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(
            lib.fresco_nuwu_batch_create(
                self.npatterns, self.npts.ctypes.data_as(_lib.c_int64_p), _lib.dptr(self._radii), self.ndim,
                self.nmodes_per_pattern.ctypes.data_as(_lib.c_int64_p), _lib.dptr(self._Sk), _lib.dptr(self._V),
                _lib.dptr(self.L), self.eps, code, float(sigma or 0.0), ctypes.byref(h),
            )
        )
        self._init_handle(h)

    def pattern_energies(self):
        out = np.empty(self.npatterns)
        _lib.check(_lib.load().fresco_pot_pattern_energies(self._h, _lib.dptr(out), self.npatterns))
        return out

    def split(self, x):
        """The per-pattern pieces of a concatenated vector."""
        return [x[self.offsets[p]:self.offsets[p + 1]] for p in range(self.npatterns)]

    def x_grad_scale(self):
        return np.tile(2 * np.pi / self.L, int(self.npts.sum()))
