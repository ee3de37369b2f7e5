"""ctypes driver of the reference's own pair potentials compiled from
source/fresco/*.hpp by `make -C oracle ref` (oracle/_ref/libref_pair.so).

TEST INFRASTRUCTURE ONLY.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libref_pair.so")
_lib = None

METHODS = {"cartesian": 0, "periodic": 1, "cells": 2}
KINDS = {"inverse_power": 0, "wca": 1}


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.ref_pair_create.restype = ctypes.c_void_p
        L.ref_pair_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp,
                                      ctypes.c_int64, ctypes.c_int, dp, ctypes.c_double]
        L.ref_pair_energy.restype = ctypes.c_double
        L.ref_pair_energy.argtypes = [ctypes.c_void_p, dp, ctypes.c_int64]
        L.ref_pair_energy_gradient.restype = ctypes.c_double
        L.ref_pair_energy_gradient.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_int64]
        L.ref_pair_destroy.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class RefPair:
    def __init__(self, kind, method, p1, eps, radii, box, ncellx_scale=1.0):
        self.radii = np.ascontiguousarray(radii, dtype=np.float64)
        self.box = np.ascontiguousarray(box, dtype=np.float64)
        self.d = len(self.box)
        self.h = lib().ref_pair_create(KINDS[kind], METHODS[method], p1, eps, _p(self.radii), len(self.radii),
                                       self.d, _p(self.box), ncellx_scale)
        if not self.h:
            raise RuntimeError("reference constructor failed")

    def get_energy(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return lib().ref_pair_energy(self.h, _p(x), x.size)

    def get_energy_gradient(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        e = lib().ref_pair_energy_gradient(self.h, _p(x), _p(g), x.size)
        return e, g

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_pair_destroy(self.h)
            self.h = None
