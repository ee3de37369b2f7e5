"""Python surface of the device optimizers.

Mirrors fresco/optimize/opt.pyx:17-93 (GradientOptimizer: one_iteration with
events, run(niter, isave, file_name), reset, stop_criterion_satisfied,
get_maxiter, get_niter, get_result), fresco/optimize/_lbfgs_cpp.pyx:43-80
(LBFGS_CPP constructor and defaults) and fresco/optimize/_modified_fire_cpp.pyx:
18-41 (ModifiedFireCPP).  The optimizer state lives on the GPU
(libfresco_hip's LBFGS / MODIFIED_FIRE); the potential is any device
potential or, through as_cpp_potential, a Python one.

Checkpoints: the reference dumps coordinates with hickle to
<file_name>_odd.hkl / _even.hkl every `isave` iterations (opt.pyx:41-54);
hickle is not available here, so the same alternation is written with
numpy.savez to <file_name>_odd.npz / _even.npz (array "coords").
"""
from __future__ import annotations

import copy
import ctypes

import numpy as np

from .. import _lib
from ..potentials.combine import as_cpp_potential
from .result import Result


class GradientOptimizer:
    _h = None

    def _finish_init(self, events):
        self.events = list(events) if events is not None else []
        self.res = None

    def _set(self, param, value):
        _lib.check(_lib.load().fresco_opt_set_param(self._h, param, float(value)))

    def one_iteration(self):
        _lib.check(_lib.load().fresco_opt_one_iteration(self._h))
        self.res = self.get_result()
        for event in self.events:
            event(coords=self.res.coords, energy=self.res.energy, rms=self.res.rms)
        return self.res

    def run(self, niter=None, isave=None, file_name="unknown"):
        lib = _lib.load()
        if not self.events:
            if niter is None and isave is None:
                _lib.check(lib.fresco_opt_run(self._h, -1))
            elif isave is None:
                _lib.check(lib.fresco_opt_run(self._h, int(niter)))
            else:
                if niter is None:
                    niter = self.get_maxiter() - self.get_niter()
                i = 0
                odd = True
                while i < niter:
                    _lib.check(lib.fresco_opt_run(self._h, int(isave)))
                    self.res = self.get_result()
                    np.savez(file_name + ("_odd.npz" if odd else "_even.npz"), coords=self.res.coords)
                    i += isave
                    odd = not odd
            self.res = self.get_result()
        else:
            if niter is None:
                niter = self.get_maxiter() - self.get_niter()
            for _ in range(int(niter)):
                if self.stop_criterion_satisfied():
                    break
                self.res = self.one_iteration()
            if self.res is None:
                self.res = self.get_result()
        return copy.deepcopy(self.res)

    def reset(self, coords):
        x = np.ascontiguousarray(np.asarray(coords, dtype=np.float64).ravel())
        _lib.check(_lib.load().fresco_opt_reset(self._h, _lib.dptr(x), len(x)))

    def stop_criterion_satisfied(self):
        v = ctypes.c_int()
        _lib.check(_lib.load().fresco_opt_stop_criterion(self._h, ctypes.byref(v)))
        return bool(v.value)

    def _state(self):
        f, rms, H0 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        it, nf, mx = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(
            _lib.load().fresco_opt_get_state(
                self._h, ctypes.byref(f), ctypes.byref(rms), ctypes.byref(it), ctypes.byref(nf), ctypes.byref(mx),
                ctypes.byref(H0),
            )
        )
        return f.value, rms.value, it.value, nf.value, mx.value, H0.value

    def get_maxiter(self):
        return self._state()[4]

    def get_niter(self):
        return self._state()[2]

    def get_result(self):
        lib = _lib.load()
        f, rms, it, nf, mx, H0 = self._state()
        x = np.empty(self._n)
        g = np.empty(self._n)
        _lib.check(lib.fresco_opt_get_x(self._h, _lib.dptr(x)))
        _lib.check(lib.fresco_opt_get_g(self._h, _lib.dptr(g)))
        res = Result()
        res.energy = f
        res.coords = x
        res.grad = g
        res.rms = rms
        res.nsteps = it
        res.nfev = nf
        res.success = self.stop_criterion_satisfied()
        return res

    def graph_stats(self):
        """(graphs recorded, iterations replayed from a graph)."""
        c, r = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.load().fresco_opt_graph_stats(self._h, ctypes.byref(c), ctypes.byref(r)))
        return int(c.value), int(r.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().fresco_opt_destroy(h)
            except Exception:
                pass
            self._h = None


class LBFGS_CPP(GradientOptimizer):
    """LBFGS with backtracking line search on the GPU (lbfgs.cpp).

    use_graph (not in the reference): replay each iteration as one recorded
    HIP graph when the potential can be recorded (NUwU, batched NUwU, UwU,
    UwNU, NUwNU);
    False launches the same kernels one by one, with identical results.
    recursion (not in the reference): "auto" (vector-free from 65536
    unknowns on when M <= 16), "two-loop" (lbfgs.cpp as written) or
    "vector-free" (the recursion on the coefficients of the memory vectors
    and g, Chen, Wang & Zhou 2014: the same step up to rounding, each memory
    vector read twice per iteration instead of about five times).
    """

    def __init__(self, x0, potential, tol=1e-5, M=4, maxstep=0.1, maxErise=1e-4, H0=0.1, iprint=-1,
                 energy=None, gradient=None, nsteps=10000, verbosity=0, events=None, logger=None,
                 rel_energy=False, use_graph=True, recursion="auto"):
        x = np.ascontiguousarray(np.asarray(x0, dtype=np.float64).ravel())
        self._n = len(x)
        self.pot = as_cpp_potential(potential, verbose=verbosity > 0, ndof=self._n)
        if logger is not None:
            print("warning c++ LBFGS is ignoring logger")
        h = ctypes.c_void_p()
        _lib.check(_lib.load().fresco_lbfgs_create(self.pot._h, _lib.dptr(x), self._n, float(tol), int(M),
                                                    ctypes.byref(h)))
        self._h = h
        self._set(_lib.OPT_H0, H0)
        self._set(_lib.OPT_MAXSTEP, maxstep)
        self._set(_lib.OPT_MAX_F_RISE, maxErise)
        self._set(_lib.OPT_MAX_ITER, nsteps)
        self._set(_lib.OPT_VERBOSITY, verbosity)
        self._set(_lib.OPT_IPRINT, iprint)
        if rel_energy:
            self._set(_lib.OPT_USE_RELATIVE_F, 1)
        self._set(_lib.OPT_USE_GRAPH, 1 if use_graph else 0)
        forms = {"auto": 0, "two-loop": 1, "vector-free": 2}
        if recursion not in forms:
            raise ValueError("recursion must be one of %s" % sorted(forms))
        self._set(_lib.OPT_LBFGS_FORM, forms[recursion])
        if energy is not None and gradient is not None:
            g = np.ascontiguousarray(np.asarray(gradient, dtype=np.float64).ravel())
            _lib.check(_lib.load().fresco_opt_set_func_gradient(self._h, float(energy), _lib.dptr(g), len(g)))
        self._finish_init(events)

    def set_H0(self, H0):
        self._set(_lib.OPT_H0, float(H0))

    def get_result(self):
        res = super().get_result()
        res["H0"] = float(self._state()[5])
        return res


class ModifiedFireCPP(GradientOptimizer):
    """Modified FIRE on the GPU (modified_fire.hpp).

    use_graph (not in the reference): keep the iteration's scalars (dt, the
    mixing, the norms, the step count) on the device and replay each
    iteration as a recorded HIP graph when the potential can be recorded;
    False runs the host-controlled iteration, with identical results.
    """

    def __init__(self, x0, potential, dtstart=0.1, dtmax=1, maxstep=0.5, Nmin=5, finc=1.1, fdec=0.5, fa=0.99,
                 astart=0.1, tol=1e-3, stepback=True, iprint=-1, energy=None, gradient=None, nsteps=10000,
                 verbosity=0, events=None, use_graph=True):
        x = np.ascontiguousarray(np.asarray(x0, dtype=np.float64).ravel())
        self._n = len(x)
        self.pot = as_cpp_potential(potential, verbose=verbosity > 0, ndof=self._n)
        h = ctypes.c_void_p()
        _lib.check(
            _lib.load().fresco_fire_create(
                self.pot._h, _lib.dptr(x), self._n, float(dtstart), float(dtmax), float(maxstep), int(Nmin),
                float(finc), float(fdec), float(fa), float(astart), float(tol), int(bool(stepback)), ctypes.byref(h),
            )
        )
        self._h = h
        self._set(_lib.OPT_MAX_ITER, nsteps)
        self._set(_lib.OPT_VERBOSITY, verbosity)
        self._set(_lib.OPT_IPRINT, iprint)
        self._set(_lib.OPT_USE_GRAPH, 1 if use_graph else 0)
        if energy is not None and gradient is not None:
            g = np.ascontiguousarray(np.asarray(gradient, dtype=np.float64).ravel())
            _lib.check(_lib.load().fresco_opt_set_func_gradient(self._h, float(energy), _lib.dptr(g), len(g)))
        self._finish_init(events)
