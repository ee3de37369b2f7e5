"""UwU: grid intensities, uniform k grid (FFT only).

Python surface of fresco/potentials/uwu.pyx:9-38 (constructor arguments, the
ifftshift of Sk / V, pickling through the constructor arguments) over the
device implementation of source/fresco/uwu.hpp.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from ._base import BasePotential, precision_code


class UwU(BasePotential):
    """phi = sum_k V_k ((|rho_k|^2/sum(x) - S0_k)/S0_k)^2 with rho the
    unnormalised +1 DFT of the intensity grid x (shape N, row-major).

    Sk and V are given centred (zero frequency in the middle) and shifted to
    FFT order with ``np.fft.ifftshift`` as the reference wrapper does
    (uwu.pyx:14-15).  The gradient is the reference's
    Re(DFT_{-1}((2/sum x) V D rho / S0)) (uwu.hpp:152-194); keyword-only
    ``exact_grad=True`` returns the exact derivative of phi instead (factor 4
    and the term from d(sum x)/dx).  ``precision``: "fp64" (default) or
    "fp32" transforms.
    """

    def __init__(self, N, K, Sk, V, L, eps=1e-6, *, precision="fp64", exact_grad=False):
        self.N = np.asarray(N).astype(np.int64).ravel()
        self.K = float(K)
        self.Sk = np.ascontiguousarray(np.fft.ifftshift(np.asarray(Sk, dtype=np.float64)).ravel())
        self.V = np.ascontiguousarray(np.fft.ifftshift(np.asarray(V, dtype=np.float64)).ravel())
        assert len(self.Sk) == len(self.V)
        self.L = np.ascontiguousarray(np.asarray(L, dtype=np.float64).ravel())
        self.ndim = len(self.L)
        self.eps = float(eps)
        self.precision = precision
        self.exact_grad = bool(exact_grad)
        if len(self.N) != self.ndim:
            raise ValueError("len(N) must equal len(L)")
        code = precision_code(precision)
        shape = np.ascontiguousarray(self.N)
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(
            lib.fresco_uwu_create(
                shape.ctypes.data_as(_lib.c_int64_p), self.ndim, _lib.dptr(self.Sk), _lib.dptr(self.V),
                len(self.V), code, int(self.exact_grad), ctypes.byref(h),
            )
        )
        self._init_handle(h)

    def x_grad_scale(self):
        if not self.exact_grad:
            raise ValueError("the reference UwU gradient is not the derivative of its energy; "
                             "construct with exact_grad=True to differentiate through it")
        return np.ones(int(np.prod(self.N)))

    def targets(self):
        shape = tuple(int(v) for v in self.N)
        return (np.fft.fftshift(self._current("Sk").reshape(shape)),
                np.fft.fftshift(self._current("V").reshape(shape)))

    def _targets_flat(self, S0, V):
        # centred grids -> FFT order, as the constructor (uwu.pyx:14-15)
        import torch

        return torch.fft.ifftshift(S0).reshape(-1), torch.fft.ifftshift(V).reshape(-1)

    def _grads_user(self, dS0, dV, like):
        import torch

        shape = tuple(int(v) for v in self.N)
        if like.dim() == 1:
            # a flat array was shifted as 1D by the constructor's ifftshift
            return torch.fft.fftshift(dS0), torch.fft.fftshift(dV)
        return torch.fft.fftshift(dS0.reshape(shape)).reshape(like.shape), torch.fft.fftshift(
            dV.reshape(shape)).reshape(like.shape)

    def __reduce__(self):
        # Sk / V are stored shifted; undo the shift so the constructor reapplies it
        shape = tuple(int(v) for v in self.N)
        Sk = np.fft.fftshift(self._current("Sk").reshape(shape))
        V = np.fft.fftshift(self._current("V").reshape(shape))
        return (_rebuild_uwu, (self.N, self.K, Sk, V, self.L, self.eps, self.precision, self.exact_grad))


def _rebuild_uwu(N, K, Sk, V, L, eps, precision, exact_grad):
    return UwU(N, K, Sk, V, L, eps, precision=precision, exact_grad=exact_grad)
