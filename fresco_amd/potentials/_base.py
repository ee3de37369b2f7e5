"""Python surface of the device potentials.

Mirrors fresco/potentials/_fresco.pyx:10-24 (BasePotential.get_energy /
get_energy_gradient) and the per-potential Cython wrappers
(fresco/potentials/nuwu.pyx:9-44 etc.): same constructor arguments, same
method names, plus the camelCase aliases the reference's PythonPotential
route calls (source/fresco/python_potential_wrapper.hpp:89,136).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from .._device import as_device_f64, device, stream_handle

PRECISIONS = {
    "fp64": _lib.PREC_FP64_PLAIN,
    "fp64-plain": _lib.PREC_FP64_PLAIN,
    "fp64-adjoint": _lib.PREC_FP64_ADJOINT,
    "fp32": _lib.PREC_FP32_ADJOINT,
    "fp32-adjoint": _lib.PREC_FP32_ADJOINT,
}


def precision_code(precision):
    if isinstance(precision, int):
        if precision not in (0, 1, 2):
            raise ValueError("unknown precision %r" % precision)
        return precision
    try:
        return PRECISIONS[str(precision).lower()]
    except KeyError:
        raise ValueError("precision must be one of %s" % sorted(PRECISIONS)) from None


class BasePotential:
    """A potential evaluated on the GPU through libfresco_hip.

    get_energy(x) -> float
    get_energy_gradient(x) -> (float, grad)

    x may be a numpy array / list (host) or a torch tensor (any device); grad
    is returned as a numpy array for host input and as a float64 device
    tensor for tensor input.
    """

    _h = None

    def _init_handle(self, handle):
        self._h = handle
        lib = _lib.load()
        n = ctypes.c_int64()
        _lib.check(lib.fresco_pot_ndof(self._h, ctypes.byref(n)))
        self._ndof = int(n.value)
        _lib.check(lib.fresco_pot_nmodes(self._h, ctypes.byref(n)))
        self._nmodes = int(n.value)

    @property
    def ndof(self):
        return self._ndof

    @property
    def nmodes(self):
        return self._nmodes

    def get_energy(self, x):
        lib = _lib.load()
        xt, _ = as_device_f64(x, self._ndof)
        e = ctypes.c_double()
        _lib.check(lib.fresco_pot_energy(self._h, ctypes.c_void_p(xt.data_ptr()), ctypes.byref(e), stream_handle()))
        return float(e.value)

    def get_energy_gradient(self, x):
        lib = _lib.load()
        xt, was_torch = as_device_f64(x, self._ndof)
        g = torch.empty_like(xt)
        e = ctypes.c_double()
        _lib.check(
            lib.fresco_pot_energy_gradient(
                self._h, ctypes.c_void_p(xt.data_ptr()), ctypes.c_void_p(g.data_ptr()), ctypes.byref(e),
                stream_handle(),
            )
        )
        if was_torch:
            return float(e.value), g
        return float(e.value), g.cpu().numpy()

    # names used by the reference's PythonPotential route
    def getEnergy(self, x):
        return self.get_energy(x)

    def getEnergyGradient(self, x):
        return self.get_energy_gradient(x)

    def energy_gradient_device(self, x, grad, stream=None):
        """Zero-copy call: x, grad are contiguous float64 device tensors."""
        lib = _lib.load()
        e = ctypes.c_double()
        s = stream_handle() if stream is None else ctypes.c_void_p(stream.cuda_stream)
        _lib.check(
            lib.fresco_pot_energy_gradient(
                self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(grad.data_ptr()), ctypes.byref(e), s
            )
        )
        return float(e.value)

    def get_rho(self):
        """rho_k of the last evaluation (complex128 numpy array, reference
        flat mode order)."""
        lib = _lib.load()
        out = torch.empty(2 * self._nmodes, dtype=torch.float64, device=device())
        _lib.check(lib.fresco_pot_get_rho(self._h, ctypes.c_void_p(out.data_ptr()), stream_handle()))
        a = out.cpu().numpy()
        return a[0::2] + 1j * a[1::2]

    def plan_info(self):
        lib = _lib.load()
        nf = (ctypes.c_int64 * 3)()
        tile = (ctypes.c_int * 3)()
        w = ctypes.c_int()
        beta = ctypes.c_double()
        _lib.check(lib.fresco_pot_plan_info(self._h, nf, ctypes.byref(w), ctypes.byref(beta), tile))
        return {"nf": list(nf), "w": int(w.value), "beta": float(beta.value), "tile": list(tile)}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().fresco_pot_destroy(h)
            except Exception:
                pass
            self._h = None
