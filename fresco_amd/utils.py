"""Analysis helpers of fresco.utils on the GPU.

export_structure (fresco/utils/utils.py:16-52) and radial_profile
(utils.py:138-152) with the reference's arguments and return values.  The
structure factor is the GPU type-1 NUFFT of the NUwU path (libfresco_hip's
fresco_structure_factor), g(r) its inverse FFT, and the shell averages the
GPU shell sums of fresco_radial_profile.

Coordinates: the reference scales r / L, shifts by -1/2 when a coordinate
exceeds 1 and multiplies by 2 pi (utils.py:19-24).  A common shift only
multiplies every rho_k by one phase, so S(k) = |rho_k|^2 / N does not depend
on it and the device map is x = 2 pi r / L.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._device import as_device_f64, device, stream_handle

__all__ = ["export_structure", "radial_profile"]


def _weights(radii, N, ndim):
    # utils.py:26-35 (volumes to the first power, normalised to sum N)
    if radii is None:
        return np.ones(N)
    radii = np.asarray(radii, dtype=np.float64).ravel()
    if ndim == 2:
        c = np.pi * radii ** 2
    elif ndim == 3:
        c = 4 / 3 * np.pi * radii ** 3
    else:
        raise ValueError("Invalid dimension")
    return c * (N / np.sum(c))


def export_structure(r, L, Nk=1001, radii=None, tol=1e-9, precision="fp64"):
    """S(k) on the Nk^d grid, its radial average, and g(r) (utils.py:16-52).

    Returns (q, S, Sk, r / Nk, g, gr) as numpy arrays: Sk and gr have shape
    (Nk,) * d with the first index along x and k = index - Nk // 2; q, S and
    r, g are their radial profiles about the centre Nk // 2.

    precision: "fp64" (the reference's double-precision finufft) or "fp32".
    Dimensions other than 2 and 3 print 'Invalid dimension' and return None,
    as the reference does.
    """
    ndim = len(L)
    L = np.asarray(L, dtype=np.float64).ravel()
    rr = np.array(r, dtype=np.float64).reshape((-1, ndim))
    N = rr.shape[0]
    if ndim not in (2, 3):
        print("Invalid dimension")
        return None
    c = _weights(radii, N, ndim)
    prec = {"fp64": 0, "fp32": 1}.get(str(precision).lower())
    if prec is None:
        raise ValueError("precision must be 'fp64' or 'fp32'")
    Nk = int(Nk)
    dev = device()
    x = torch.as_tensor(rr.ravel(), dtype=torch.float64).to(dev)
    cw = torch.as_tensor(c, dtype=torch.float64).to(dev)
    shape = (Nk,) * ndim
    Sk = torch.empty(shape, dtype=torch.float64, device=dev)
    gr = torch.empty(shape, dtype=torch.float64, device=dev)
    _lib.check(
        _lib.load().fresco_structure_factor(
            ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(cw.data_ptr()), N, ndim, _lib.dptr(L), Nk,
            float(tol), prec, ctypes.c_void_p(Sk.data_ptr()), ctypes.c_void_p(gr.data_ptr()), stream_handle(),
        )
    )
    center = int(Nk / 2)
    q, S = radial_profile(Sk, center=[center] * ndim)
    rq, g = radial_profile(gr, center=[center] * ndim)
    return q, S, Sk.cpu().numpy(), rq / Nk, g, gr.cpu().numpy()


def radial_profile(data, center=None):
    """Shell averages of a 1-3 dimensional grid (utils.py:138-152).

    Shell b holds the cells with floor(|index - center|) = b; cells beyond
    R = shape[0] // 2 are dropped.  Returns (shells present, averages) like
    the reference: the last shell present and its average are left out.
    data: numpy array or torch tensor (real).
    """
    if isinstance(data, torch.Tensor):
        if data.is_complex():
            raise TypeError("radial_profile needs real data")
        shape = tuple(data.shape)
    else:
        data = np.asarray(data)
        if np.iscomplexobj(data):
            raise TypeError("radial_profile needs real data")
        shape = data.shape
    ndim = len(shape)
    if ndim < 1 or ndim > 3:
        raise ValueError("radial_profile takes 1 to 3 dimensional grids")
    if center is None:
        center = np.array(shape) / 2
    center = np.asarray(center, dtype=np.float64).ravel()
    if center.size != ndim:
        raise ValueError("center needs %d entries" % ndim)
    R = int(shape[0] / 2)
    dev_data, _ = as_device_f64(data)
    nb = R + 2
    sums = torch.empty(nb, dtype=torch.float64, device=dev_data.device)
    counts = torch.empty(nb, dtype=torch.int64, device=dev_data.device)
    shp = np.asarray(shape, dtype=np.int64)
    _lib.check(
        _lib.load().fresco_radial_profile(
            ctypes.c_void_p(dev_data.data_ptr()), ndim, shp.ctypes.data_as(_lib.c_int64_p), _lib.dptr(center),
            ctypes.c_void_p(sums.data_ptr()), ctypes.c_void_p(counts.data_ptr()), stream_handle(),
        )
    )
    sums = sums.cpu().numpy()
    counts = counts.cpu().numpy()
    present = np.nonzero(counts)[0]
    last = int(present[-1]) + 1 if present.size else 0
    tbin, nr = sums[:last], counts[:last]
    with np.errstate(divide="ignore", invalid="ignore"):
        prof = tbin / nr
    return present[:-1], prof[:-1]
