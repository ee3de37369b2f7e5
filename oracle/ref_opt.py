"""ctypes driver of the reference's own LBFGS / modified FIRE compiled from its
sources into oracle/_ref/libref_opt.so (oracle/ref_shim.cpp).

TEST INFRASTRUCTURE ONLY.  Absent (None) when the library was not built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libref_opt.so")
CB = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                      ctypes.c_int64, ctypes.c_void_p)
_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.ref_lbfgs_create.restype = ctypes.c_void_p
        L.ref_lbfgs_create.argtypes = [CB, ctypes.c_void_p, dp, ctypes.c_int64, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int]
        L.ref_fire_create.restype = ctypes.c_void_p
        L.ref_fire_create.argtypes = [CB, ctypes.c_void_p, dp, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int]
        L.ref_opt_run.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ref_opt_one_iteration.argtypes = [ctypes.c_void_p]
        L.ref_opt_stop.argtypes = [ctypes.c_void_p]
        L.ref_opt_stop.restype = ctypes.c_int
        L.ref_opt_state.argtypes = [ctypes.c_void_p, dp, dp, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int), dp, dp, dp]
        L.ref_opt_reset.argtypes = [ctypes.c_void_p, dp, ctypes.c_int64]
        L.ref_opt_destroy.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


class _Opt:
    def __init__(self, potential, n):
        self.n = n
        self.pot = potential

        def cb(xp, gp, nn, user):
            x = np.ctypeslib.as_array(xp, shape=(nn,)).copy()
            e, g = potential.get_energy_gradient(x)
            np.ctypeslib.as_array(gp, shape=(nn,))[:] = np.asarray(g, dtype=np.float64)
            return float(e)

        self._cb = CB(cb)
        self.h = None

    def run(self, niter=-1):
        lib().ref_opt_run(self.h, niter)

    def one_iteration(self):
        lib().ref_opt_one_iteration(self.h)

    def stop_criterion_satisfied(self):
        return bool(lib().ref_opt_stop(self.h))

    def state(self):
        f, rms, H0 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        it, nf = ctypes.c_int(), ctypes.c_int()
        x = np.zeros(self.n)
        g = np.zeros(self.n)
        dp = ctypes.POINTER(ctypes.c_double)
        lib().ref_opt_state(self.h, ctypes.byref(f), ctypes.byref(rms), ctypes.byref(it), ctypes.byref(nf),
                            ctypes.byref(H0), x.ctypes.data_as(dp), g.ctypes.data_as(dp))
        return {"f": f.value, "rms": rms.value, "niter": it.value, "nfev": nf.value, "H0": H0.value, "x": x, "g": g}

    def reset(self, x0):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        lib().ref_opt_reset(self.h, x0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(x0))

    def __del__(self):
        if self.h:
            lib().ref_opt_destroy(self.h)
            self.h = None


class LBFGS(_Opt):
    def __init__(self, x0, potential, tol=1e-4, M=4, maxstep=0.1, max_f_rise=1e-4, H0=0.1, rel=False,
                 max_iter=1000):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        super().__init__(potential, len(x0))
        self.h = lib().ref_lbfgs_create(self._cb, None, x0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        len(x0), tol, M, maxstep, max_f_rise, H0, int(rel), max_iter)


class ModifiedFire(_Opt):
    def __init__(self, x0, potential, dtstart=0.1, dtmax=1.0, maxstep=0.5, Nmin=5, finc=1.1, fdec=0.5, fa=0.99,
                 astart=0.1, tol=1e-4, stepback=True, max_iter=1000):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        super().__init__(potential, len(x0))
        self.h = lib().ref_fire_create(self._cb, None, x0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       len(x0), dtstart, dtmax, maxstep, Nmin, finc, fdec, fa, astart, tol,
                                       int(stepback), max_iter)


# --- the reference's optimizers over a GPU potential -------------------------
# oracle/ref_hip_lbfgs.cpp: source/lbfgs.cpp / modified_fire.* with a
# libfresco_hip potential behind include/fresco_hip_potential.hpp
# (fresco::HipPotential), built into oracle/_ref/libref_hip_opt.so.

HIP_LIB_PATH = os.path.join(_HERE, "_ref", "libref_hip_opt.so")
_hip_lib = None


def hip_available():
    return os.path.exists(HIP_LIB_PATH)


def hip_lib():
    global _hip_lib
    if _hip_lib is None:
        L = ctypes.CDLL(HIP_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.ref_hip_lbfgs_create.restype = vp
        L.ref_hip_lbfgs_create.argtypes = [vp, dp, ctypes.c_int64, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.ref_hip_fire_create.restype = vp
        L.ref_hip_fire_create.argtypes = [vp, dp, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.ref_hip_one_iteration.argtypes = [vp]
        L.ref_hip_run.argtypes = [vp, ctypes.c_int]
        L.ref_hip_state.argtypes = [vp, dp, dp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), dp, dp]
        L.ref_hip_stop.argtypes = [vp]
        L.ref_hip_stop.restype = ctypes.c_int
        L.ref_hip_destroy.argtypes = [vp]
        _hip_lib = L
    return _hip_lib


class _HipOpt:
    """pot_handle: the fresco_pot* (ctypes.c_void_p) of a potential that stays
    alive, and owned, on the Python side."""

    def __init__(self, n):
        self.n = n
        self.h = None

    def one_iteration(self):
        if hip_lib().ref_hip_one_iteration(self.h) != 0:
            raise RuntimeError("reference optimizer raised")

    def run(self, niter):
        if hip_lib().ref_hip_run(self.h, niter) != 0:
            raise RuntimeError("reference optimizer raised")

    def stop_criterion_satisfied(self):
        return bool(hip_lib().ref_hip_stop(self.h))

    def state(self):
        f, rms = ctypes.c_double(), ctypes.c_double()
        it, nf = ctypes.c_int(), ctypes.c_int()
        x = np.zeros(self.n)
        g = np.zeros(self.n)
        dp = ctypes.POINTER(ctypes.c_double)
        hip_lib().ref_hip_state(self.h, ctypes.byref(f), ctypes.byref(rms), ctypes.byref(it), ctypes.byref(nf),
                                x.ctypes.data_as(dp), g.ctypes.data_as(dp))
        return {"f": f.value, "rms": rms.value, "niter": it.value, "nfev": nf.value, "x": x, "g": g}

    def __del__(self):
        if self.h:
            hip_lib().ref_hip_destroy(self.h)
            self.h = None


class HipLBFGS(_HipOpt):
    def __init__(self, pot_handle, x0, tol=1e-4, M=4, maxstep=0.1, max_f_rise=1e-4, H0=0.1, max_iter=1000):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        super().__init__(len(x0))
        self.h = hip_lib().ref_hip_lbfgs_create(pot_handle, x0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                len(x0), tol, M, maxstep, max_f_rise, H0, max_iter)
        if not self.h:
            raise RuntimeError("reference LBFGS over the GPU potential could not be built")


class HipModifiedFire(_HipOpt):
    def __init__(self, pot_handle, x0, dtstart=0.1, dtmax=1.0, maxstep=0.5, Nmin=5, finc=1.1, fdec=0.5, fa=0.99,
                 astart=0.1, tol=1e-4, stepback=True, max_iter=1000):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        super().__init__(len(x0))
        self.h = hip_lib().ref_hip_fire_create(pot_handle, x0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                               len(x0), dtstart, dtmax, maxstep, Nmin, finc, fdec, fa, astart, tol,
                                               int(stepback), max_iter)
        if not self.h:
            raise RuntimeError("reference FIRE over the GPU potential could not be built")
