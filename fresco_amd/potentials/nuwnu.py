"""NUwNU: non-uniform points, non-uniform k vectors.

Python surface of fresco/potentials/nuwnu.pyx:9-38 over the device
implementation of source/fresco/nuwnu.hpp (evaluated here as exact sums with
float64 phases).
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from ._base import BasePotential, precision_code


class NUwNU(BasePotential):
    """phi = sum_q V_q ((|rho_q|^2/N - S0_q)/S0_q)^2 with
    rho_q = sum_j c_j exp(+i K_q . x_j), x = (p - round(p/L) + L) 2 pi / L
    and K used as given (nuwnu.hpp:126-182).  The gradient is
    Re(sum_q K_q,a (-4i/N) V D rho / S0 exp(-i K_q . x_j)) with no weight
    factor (nuwnu.hpp:184-290).  Nk = len(K) / d.

    As in the reference, rho and the gradient come from type-3 NUFFTs at
    tolerance ``eps``.  Keyword-only extensions: ``precision`` ("fp64"
    (default) or "fp32" grids), ``method`` ("nufft" (default) or "direct":
    exact sums with float64 phases, O(N Nk)) and ``sigma`` (upsampling).
    """

    def __init__(self, radii, K, Sk, V, L, eps=1e-6, *, precision="fp64", method="nufft", sigma=None):
        self.radii = np.ascontiguousarray(np.asarray(radii, dtype=np.float64).ravel())
        self.N = len(self.radii)
        self.K = np.ascontiguousarray(np.asarray(K, dtype=np.float64).ravel())
        self.Sk = np.ascontiguousarray(np.asarray(Sk, dtype=np.float64).ravel())
        self.V = np.ascontiguousarray(np.asarray(V, dtype=np.float64).ravel())
        assert len(self.Sk) == len(self.V)
        self.L = np.ascontiguousarray(np.asarray(L, dtype=np.float64).ravel())
        self.ndim = len(self.L)
        self.eps = float(eps)
        self.precision = precision
        self.method = method
        self.sigma = sigma
        code = precision_code(precision)
        methods = {"nufft": 0, "direct": 1}
        if method not in methods:
            raise ValueError("method must be 'nufft' or 'direct'")
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(
            lib.fresco_nuwnu_create(
                _lib.dptr(self.radii), self.N, self.ndim, _lib.dptr(self.K), len(self.K), _lib.dptr(self.Sk),
                _lib.dptr(self.V), len(self.V), _lib.dptr(self.L), self.eps, code, methods[method],
                float(sigma or 0.0), ctypes.byref(h),
            )
        )
        self._init_handle(h)

    def x_grad_scale(self):
        return np.tile(2 * np.pi / self.L, self.N)

    def __reduce__(self):
        return (_rebuild_nuwnu, (self.radii, self.K, self._current("Sk"), self._current("V"), self.L, self.eps,
                                 self.precision, self.method, self.sigma))


def _rebuild_nuwnu(radii, K, Sk, V, L, eps, precision, method, sigma):
    return NUwNU(radii, K, Sk, V, L, eps, precision=precision, method=method, sigma=sigma)
