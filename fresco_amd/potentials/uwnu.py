"""UwNU: grid values as NUFFT modes, non-uniform k vectors.

Python surface of fresco/potentials/uwnu.pyx:9-38 over the device
implementation of source/fresco/uwnu.hpp.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from ._base import BasePotential, precision_code


class UwNU(BasePotential):
    """phi = sum_q V_q ((|rho_q|^2 - S0_q)/S0_q)^2 with
    rho_q = sum_m x_m exp(+i k_q . m), k_q = 2 pi K_q / L and m the integer
    mode index of grid cell (i0, i1, i2), m_a = i_a - N_a // 2, the first
    dimension fastest (FINUFFT mode order, uwnu.hpp:44-82).  The gradient is
    Re(type-1 (-1) of 4 V D rho / S0) / (2 pi) (uwnu.hpp:117-158).

    Keyword-only extensions: ``precision`` ("fp64" default, "fp32") and
    ``sigma`` (fine-grid upsampling factor, default 2).
    """

    def __init__(self, N, K, Sk, V, L, eps=1e-6, *, precision="fp64", sigma=None):
        self.N = np.asarray(N).astype(np.int64).ravel()
        self.K = np.ascontiguousarray(np.asarray(K, dtype=np.float64).ravel())
        self.Sk = np.ascontiguousarray(np.asarray(Sk, dtype=np.float64).ravel())
        self.V = np.ascontiguousarray(np.asarray(V, dtype=np.float64).ravel())
        assert len(self.Sk) == len(self.V)
        self.L = np.ascontiguousarray(np.asarray(L, dtype=np.float64).ravel())
        self.ndim = len(self.L)
        self.eps = float(eps)
        self.precision = precision
        self.sigma = sigma
        if len(self.N) != self.ndim:
            raise ValueError("len(N) must equal len(L)")
        code = precision_code(precision)
        shape = np.ascontiguousarray(self.N)
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(
            lib.fresco_uwnu_create(
                shape.ctypes.data_as(_lib.c_int64_p), self.ndim, _lib.dptr(self.K), len(self.K),
                _lib.dptr(self.Sk), _lib.dptr(self.V), len(self.V), _lib.dptr(self.L), self.eps, code,
                float(sigma or 0.0), ctypes.byref(h),
            )
        )
        self._init_handle(h)

    def x_grad_scale(self):
        # uwnu.hpp:155 divides the derivative by 2 pi
        return np.full(int(np.prod(self.N)), 2 * np.pi)

    def __reduce__(self):
        return (_rebuild_uwnu, (self.N, self.K, self._current("Sk"), self._current("V"), self.L, self.eps, self.precision, self.sigma))


def _rebuild_uwnu(N, K, Sk, V, L, eps, precision, sigma):
    return UwNU(N, K, Sk, V, L, eps, precision=precision, sigma=sigma)
