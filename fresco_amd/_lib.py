"""ctypes binding of libfresco_hip.so (the C ABI declared in include/fresco_hip.h).

The library is the only compute path: if it is missing, or no GPU is visible,
every potential fails loudly here; nothing falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfresco_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "fresco_hip.h")

FRESCO_OK = 0
FRESCO_E_INVALID = 1
FRESCO_E_RUNTIME = 2
FRESCO_E_HIP = 3

PREC_FP64_PLAIN = 0
PREC_FP32_ADJOINT = 1
PREC_FP64_ADJOINT = 2

_lock = threading.Lock()
_lib = None

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_void_pp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); every entry point of include/fresco_hip.h
SIGNATURES = {
    "fresco_last_error": (ctypes.c_char_p, []),
    "fresco_version": (ctypes.c_char_p, []),
    "fresco_set_fault": (ctypes.c_int, [ctypes.c_int]),
    "fresco_nuwu_create": (
        ctypes.c_int,
        [c_double_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double, c_double_p, c_double_p,
         ctypes.c_int64, c_double_p, ctypes.c_double, ctypes.c_int, ctypes.c_double, c_void_pp],
    ),
    "fresco_nuwu_batch_create": (
        ctypes.c_int,
        [ctypes.c_int, c_int64_p, c_double_p, ctypes.c_int, c_int64_p, c_double_p, c_double_p, c_double_p,
         ctypes.c_double, ctypes.c_int, ctypes.c_double, c_void_pp],
    ),
    "fresco_pot_pattern_energies": (ctypes.c_int, [ctypes.c_void_p, c_double_p, ctypes.c_int]),
    "fresco_pair_create": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.c_double, ctypes.c_double, c_double_p, ctypes.c_int64, ctypes.c_int, c_double_p,
         ctypes.c_int, ctypes.c_int, ctypes.c_double, c_void_pp],
    ),
    "fresco_uwu_create": (
        ctypes.c_int,
        [c_int64_p, ctypes.c_int, c_double_p, c_double_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, c_void_pp],
    ),
    "fresco_uwnu_create": (
        ctypes.c_int,
        [c_int64_p, ctypes.c_int, c_double_p, ctypes.c_int64, c_double_p, c_double_p, ctypes.c_int64,
         c_double_p, ctypes.c_double, ctypes.c_int, ctypes.c_double, c_void_pp],
    ),
    "fresco_nuwnu_create": (
        ctypes.c_int,
        [c_double_p, ctypes.c_int64, ctypes.c_int, c_double_p, ctypes.c_int64, c_double_p, c_double_p,
         ctypes.c_int64, c_double_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, c_void_pp],
    ),
    "fresco_pot_energy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_double_p, ctypes.c_void_p]),
    "fresco_pot_energy_gradient": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_double_p, ctypes.c_void_p],
    ),
    "fresco_pot_energy_gradient_stream": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_double_p, ctypes.c_void_p],
    ),
    "fresco_pot_get_rho": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fresco_pot_backward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fresco_pot_energy_keep": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
         ctypes.c_void_p],
    ),
    "fresco_pot_gradient_kept": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p],
    ),
    "fresco_pot_set_targets": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fresco_pot_spread_config": (ctypes.c_int, [ctypes.c_void_p, c_int_p, c_int_p]),
    "fresco_pot_spread_source": (ctypes.c_int, [ctypes.c_void_p, c_int_p]),
    "fresco_pot_ndof": (ctypes.c_int, [ctypes.c_void_p, c_int64_p]),
    "fresco_pot_nmodes": (ctypes.c_int, [ctypes.c_void_p, c_int64_p]),
    "fresco_pot_plan_info": (
        ctypes.c_int,
        [ctypes.c_void_p, c_int64_p, c_int_p, c_double_p, c_int_p],
    ),
    "fresco_pot_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fresco_combined_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, c_void_pp]),
    "fresco_callback_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, c_void_pp]),
    "fresco_lbfgs_create": (
        ctypes.c_int,
        [ctypes.c_void_p, c_double_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int, c_void_pp],
    ),
    "fresco_fire_create": (
        ctypes.c_int,
        [ctypes.c_void_p, c_double_p, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
         ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
         ctypes.c_int, c_void_pp],
    ),
    "fresco_opt_set_param": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]),
    "fresco_opt_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fresco_opt_one_iteration": (ctypes.c_int, [ctypes.c_void_p]),
    "fresco_opt_reset": (ctypes.c_int, [ctypes.c_void_p, c_double_p, ctypes.c_int64]),
    "fresco_opt_set_func_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, c_double_p, ctypes.c_int64]),
    "fresco_opt_stop_criterion": (ctypes.c_int, [ctypes.c_void_p, c_int_p]),
    "fresco_opt_get_state": (
        ctypes.c_int,
        [ctypes.c_void_p, c_double_p, c_double_p, c_int_p, c_int_p, c_int_p, c_double_p],
    ),
    "fresco_opt_get_x": (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    "fresco_opt_graph_stats": (ctypes.c_int, [ctypes.c_void_p, c_int64_p, c_int64_p]),
    "fresco_structure_factor": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, c_double_p, ctypes.c_int64,
         ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "fresco_radial_profile": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int, c_int64_p, c_double_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "fresco_radial_profile_noalloc": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int, c_int64_p, c_double_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "fresco_opt_get_g": (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    "fresco_opt_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fresco_prof_enable": (ctypes.c_int, [ctypes.c_int]),
    "fresco_prof_enable_stages": (ctypes.c_int, [ctypes.c_uint]),
    "fresco_prof_sample_period": (ctypes.c_int, [ctypes.c_int]),
    "fresco_prof_reset": (ctypes.c_int, []),
    "fresco_prof_get": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double), c_int64_p],
    ),
}

N_STAGES = 8

OPT_TOL, OPT_MAXSTEP, OPT_MAX_ITER, OPT_IPRINT, OPT_VERBOSITY, OPT_H0, OPT_MAX_F_RISE, OPT_USE_RELATIVE_F = range(8)
OPT_USE_GRAPH = 8
OPT_LBFGS_FORM = 9

EG_CALLBACK = ctypes.CFUNCTYPE(ctypes.c_int, c_double_p, c_double_p, c_double_p, ctypes.c_void_p)


def prof_enable(on=True):
    check(load().fresco_prof_enable(1 if on else 0))


def prof_enable_stages(names):
    """Time only the named stages (e.g. ["spread"])."""
    lib = load()
    mask = 0
    for st in range(N_STAGES):
        name = ctypes.c_char_p()
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        check(lib.fresco_prof_get(st, ctypes.byref(name), ctypes.byref(ms), ctypes.byref(cnt)))
        if name.value.decode() in names:
            mask |= 1 << st
    check(lib.fresco_prof_enable_stages(mask))


def prof_sample_period(period):
    """Time one launch in `period` of each enabled stage (1: every launch)."""
    check(load().fresco_prof_sample_period(int(period)))


def prof_reset():
    check(load().fresco_prof_reset())


def prof_read():
    """{stage: (total_ms, count)} of the HIP-event stage profiler."""
    lib = load()
    out = {}
    for st in range(N_STAGES):
        name = ctypes.c_char_p()
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        check(lib.fresco_prof_get(st, ctypes.byref(name), ctypes.byref(ms), ctypes.byref(cnt)))
        out[name.value.decode()] = (float(ms.value), int(cnt.value))
    return out


class FrescoError(RuntimeError):
    """Raised for HIP / rocFFT failures reported by the library."""


def load():
    """Load libfresco_hip.so (once) and declare the C signatures."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libfresco_hip.so not found at %s; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C fresco_amd/csrc`" % LIB_PATH
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int):
    """Map a status code to the Python exception the reference's Cython
    `except +` would raise (std::invalid_argument -> ValueError)."""
    if rc == FRESCO_OK:
        return
    msg = load().fresco_last_error().decode(errors="replace")
    if rc == FRESCO_E_INVALID:
        raise ValueError(msg)
    if rc == FRESCO_E_HIP:
        raise FrescoError(msg)
    raise RuntimeError(msg)


def dptr(arr):
    """ctypes double* for a C-contiguous float64 numpy array."""
    return arr.ctypes.data_as(c_double_p)
