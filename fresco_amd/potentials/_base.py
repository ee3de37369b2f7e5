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
        self._last_x = xt  # get_rho / target_grads may re-read the last coordinates
        e = ctypes.c_double()
        _lib.check(lib.fresco_pot_energy(self._h, ctypes.c_void_p(xt.data_ptr()), ctypes.byref(e), stream_handle()))
        return float(e.value)

    def get_energy_gradient(self, x):
        lib = _lib.load()
        xt, was_torch = as_device_f64(x, self._ndof)
        self._last_x = xt
        g = torch.empty_like(xt)
        e = ctypes.c_double()
        # stream-ordered: the gradient is completed by work on the current
        # stream, which the tensor's uses (and .cpu()) follow
        _lib.check(
            lib.fresco_pot_energy_gradient_stream(
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
        """Zero-copy call: x, grad are contiguous float64 device tensors.  The
        energy is returned as soon as it is known; grad is complete for work
        ordered after the call on the stream used (the current stream, or
        `stream`), as for any asynchronous torch operation."""
        lib = _lib.load()
        self._last_x = x
        e = ctypes.c_double()
        s = stream_handle() if stream is None else ctypes.c_void_p(stream.cuda_stream)
        _lib.check(
            lib.fresco_pot_energy_gradient_stream(
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

    # --- targets and their derivatives (differentiable S(k) shaping) -------

    def _targets_flat(self, S0, V):
        """(S0, V) in the flat mode order the library holds them in."""
        return S0.reshape(-1), V.reshape(-1)

    def _grads_user(self, dS0, dV, like):
        """Flat library-order target gradients -> the layout of ``like``
        (entries beyond the potential's modes, which it never reads, get 0)."""
        n = like.numel()
        if n != dS0.numel():
            dS0 = torch.cat([dS0, dS0.new_zeros(n - dS0.numel())])
            dV = torch.cat([dV, dV.new_zeros(n - dV.numel())])
        return dS0.reshape(like.shape), dV.reshape(like.shape)

    def x_grad_scale(self):
        """Factor turning the returned gradient into d phi / d x for the
        coordinates passed in (the reference returns d phi / d(2 pi p / L),
        nuwu.hpp:294-314)."""
        raise NotImplementedError

    def set_targets(self, S0, V=None):
        """Replace S0(k) and, unless V is None, V(k) (same layout as the
        constructor's Sk / V; numpy arrays or tensors).  With V None the
        weights stay and the update needs no host synchronisation."""
        S0 = torch.as_tensor(S0, dtype=torch.float64)
        keep_v = V is None
        if keep_v:
            user = getattr(self, "_user_targets", None)
            Vu = user[1] if user is not None else torch.as_tensor(self.V, dtype=torch.float64)
        else:
            Vu = torch.as_tensor(V, dtype=torch.float64)
        s, v = self._targets_flat(S0, Vu)
        s = s[: self._nmodes].to(device=device(), dtype=torch.float64).contiguous()
        if s.numel() != self._nmodes:
            raise ValueError("S0 / V must hold %d modes" % self._nmodes)
        if keep_v and getattr(self, "_dev_targets", None) is not None:
            v = self._dev_targets[1]
        else:
            v = v[: self._nmodes].to(device=device(), dtype=torch.float64).contiguous()
            if v.numel() != self._nmodes:
                raise ValueError("S0 / V must hold %d modes" % self._nmodes)
        _lib.check(
            _lib.load().fresco_pot_set_targets(
                self._h, ctypes.c_void_p(s.data_ptr()), ctypes.c_void_p(None if keep_v else v.data_ptr()),
                stream_handle(),
            )
        )
        self._dev_targets = (s, v)
        self._user_targets = (S0, Vu)
        self._ag_targets = None  # the autograd function's record of what it last set

    def targets(self):
        """Current (S0, V) in the constructor's layout (numpy)."""
        return self._current("Sk"), self._current("V")

    def _current(self, name):
        """Constructor-layout Sk / V as numpy, including set_targets updates."""
        user = getattr(self, "_user_targets", None)
        if user is None:
            return getattr(self, name)
        return user[0 if name == "Sk" else 1].detach().cpu().numpy()

    # --- split evaluation (the autograd function's forward / backward) -----

    def energy_keep(self, x):
        """Energy at x (a float64 device tensor) keeping what the gradient
        needs: (token, E) with a nonzero token naming this evaluation, or
        (0, None) for potentials without the split."""
        e = ctypes.c_double()
        token = ctypes.c_uint64()
        self._last_x = x
        _lib.check(
            _lib.load().fresco_pot_energy_keep(
                self._h, ctypes.c_void_p(x.data_ptr()), ctypes.byref(e), ctypes.byref(token), stream_handle()
            )
        )
        return (int(token.value), float(e.value)) if token.value else (0, None)

    def gradient_kept(self, token, grad=None):
        """Gradient of the kept evaluation ``token`` into grad (a float64
        device tensor of ndof entries; None: only check that it is kept).
        False when the potential no longer keeps that evaluation."""
        ok = ctypes.c_int()
        _lib.check(
            _lib.load().fresco_pot_gradient_kept(
                self._h, ctypes.c_uint64(token), ctypes.c_void_p(None if grad is None else grad.data_ptr()),
                ctypes.byref(ok), stream_handle(),
            )
        )
        return bool(ok.value)

    def target_grads(self):
        """(d phi/d S0, d phi/d V) at the last evaluation, float64 device
        tensors in the library's flat mode order."""
        dS0 = torch.empty(self._nmodes, dtype=torch.float64, device=device())
        dV = torch.empty_like(dS0)
        _lib.check(
            _lib.load().fresco_pot_backward(
                self._h, ctypes.c_void_p(dS0.data_ptr()), ctypes.c_void_p(dV.data_ptr()), stream_handle()
            )
        )
        return dS0, dV

    def plan_info(self):
        lib = _lib.load()
        nf = (ctypes.c_int64 * 3)()
        tile = (ctypes.c_int * 3)()
        w = ctypes.c_int()
        beta = ctypes.c_double()
        _lib.check(lib.fresco_pot_plan_info(self._h, nf, ctypes.byref(w), ctypes.byref(beta), tile))
        mode = ctypes.c_int()
        parts = ctypes.c_int()
        _lib.check(lib.fresco_pot_spread_config(self._h, ctypes.byref(mode), ctypes.byref(parts)))
        src = ctypes.c_int()
        _lib.check(lib.fresco_pot_spread_source(self._h, ctypes.byref(src)))
        return {"nf": list(nf), "w": int(w.value), "beta": float(beta.value), "tile": list(tile),
                "spreader": {0: "tiles", 1: "halo", 2: "halo4", 3: "halom", 4: "halor"}.get(mode.value),
                "spread_parts": int(parts.value),
                "spread_source": {0: "fixed", 1: "tuning table", 2: "autotuned", 3: "pinned"}.get(src.value)}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().fresco_pot_destroy(h)
            except Exception:
                pass
            self._h = None
