"""NUwU: non-uniform points, uniform k grid.

Python surface of fresco/potentials/nuwu.pyx:9-44 (constructor arguments,
ravel of Sk / V, public attributes, pickling through the constructor
arguments) over the device implementation of source/fresco/nuwu.hpp.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from ._base import BasePotential, precision_code


class NUwU(BasePotential):
    """Structure-factor loss phi = sum_k V_k ((|rho_k|^2/N - S0_k)/S0_k)^2 of
    N points against a (2K'+1)^d grid of integer k.

    Parameters follow the reference: radii[N], K (stored, unused), Sk and V of
    size (2K'+1)^d (any shape, C-order ravel, x fastest in the flat order), box
    L[d] and NUFFT tolerance eps.  Keyword-only extensions: ``precision``
    ("fp64" = FP64 with d type-2 transforms as in the reference (default),
    "fp64-adjoint", "fp32" = FP32 grids with the one-transform kernel-derivative
    gradient) and ``sigma`` (fine-grid upsampling factor, default 2).

    The gradient is d phi / d x with x = 2 pi p / L (nuwu.hpp:294-314).
    """

    def __init__(self, radii, K, Sk, V, L, eps=1e-6, *, precision="fp64", sigma=None):
        self.radii = np.ascontiguousarray(np.asarray(radii, dtype=np.float64).ravel())
        self.N = len(self.radii)
        self.K = float(K)
        self.Sk = np.ascontiguousarray(np.asarray(Sk, dtype=np.float64).ravel())
        self.V = np.ascontiguousarray(np.asarray(V, dtype=np.float64).ravel())
        assert len(self.Sk) == len(self.V)
        self.L = np.ascontiguousarray(np.asarray(L, dtype=np.float64).ravel())
        self.ndim = len(self.L)
        self.eps = float(eps)
        self.precision = precision
        self.sigma = sigma
        code = precision_code(precision)
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(
            lib.fresco_nuwu_create(
                _lib.dptr(self.radii), self.N, self.ndim, self.K, _lib.dptr(self.Sk), _lib.dptr(self.V),
                len(self.V), _lib.dptr(self.L), self.eps, code, float(sigma or 0.0), ctypes.byref(h),
            )
        )
        self._init_handle(h)

    def x_grad_scale(self):
        return np.tile(2 * np.pi / self.L, self.N)

    def __reduce__(self):
        return (
            _rebuild_nuwu,
            (self.radii, self.K, self._current("Sk"), self._current("V"), self.L, self.eps, self.precision, self.sigma),
        )


def _rebuild_nuwu(radii, K, Sk, V, L, eps, precision, sigma):
    return NUwU(radii, K, Sk, V, L, eps, precision=precision, sigma=sigma)
