"""Oracle of the analysis helpers export_structure and radial_profile.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

export_structure (fresco/utils/utils.py:16-52) with the finufft type-1 call
replaced by the exact sum it approximates, and radial_profile
(utils.py:138-152) restated in numpy.

Conventions taken from the Python finufft the reference calls (not in the
reference tree; finufft 2.x, Barnett, Magland & af Klinteberg 2019):
nufft{2,3}d1(x, y[, z], c, n_modes=(Nk,)*d) returns f[i1, i2(, i3)] =
sum_j c_j exp(+i (k1 x_j + k2 y_j (+ k3 z_j))), k_a = i_a - Nk // 2 (mode
order 0), first array index along x.  S = |f|^2 / N is the same for either
exponent sign because c is real.

Parity pinning (tests/test_oracle_structure.py):
  radial_profile   brute-force Python loops over every cell on tiny 1D / 2D /
                   3D grids (integer, half-integer and arbitrary centres);
                   constant data gives a constant profile; shell counts of a
                   2D grid about its centre match the lattice-point counts.
  structure_factor square / cubic lattices: S = N on the reciprocal lattice
                   (k = 0 mod m), 0 elsewhere; one point: S = 1; two points:
                   S = 1 + cos(k.dx); the first-index-along-x layout through
                   a pattern that is not symmetric under exchanging axes.
  g(r)             lattice closed form: g = (V/N) (N (Nk/m)^d [r = 0 mod
                   Nk/m] - Nk^d [r = 0]) at the centred index.
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2.0 * np.pi


def weights(radii, N, ndim):
    """utils.py:26-35."""
    if radii is None:
        return np.ones(N)
    radii = np.asarray(radii, dtype=np.float64)
    if ndim == 2:
        c = np.pi * radii ** 2
    elif ndim == 3:
        c = 4 / 3 * np.pi * radii ** 3
    else:
        raise ValueError("Invalid dimension")
    return c * (N / np.sum(c))


def mode_axis(Nk):
    return np.arange(Nk) - Nk // 2


def structure_factor(r, L, Nk, radii=None):
    """|f|^2 / N on the (Nk,)*d grid, f the exact type-1 sum of
    utils.py:38, 43 at x = 2 pi r / L (the common shift of utils.py:22-23
    only changes the phase of every f_k)."""
    L = np.asarray(L, dtype=np.float64)
    ndim = len(L)
    rr = np.asarray(r, dtype=np.float64).reshape(-1, ndim) / L * TWO_PI
    N = rr.shape[0]
    c = weights(radii, N, ndim)
    ks = mode_axis(Nk)
    # separable phases e^{i k x_j} per axis: [N, Nk]
    ph = [np.exp(1j * np.outer(rr[:, a], ks)) for a in range(ndim)]
    if ndim == 2:
        f = np.einsum("j,ja,jb->ab", c, ph[0], ph[1])
    else:
        f = np.einsum("j,ja,jb,jc->abc", c, ph[0], ph[1], ph[2])
    return np.abs(f) ** 2 / N


def pair_correlation(Sk, L, N):
    """utils.py:40-41, 45-46."""
    Nk = Sk.shape[0]
    ndim = Sk.ndim
    gr = np.fft.ifftn(np.fft.ifftshift(Sk) - 1.0, (Nk,) * ndim)
    return np.real(np.fft.fftshift(gr)) * Nk ** ndim * np.prod(L) / N


def radial_profile(data, center=None):
    """utils.py:138-152."""
    data = np.asarray(data)
    ndim = len(data.shape)
    R = int(data.shape[0] / 2)
    if center is None:
        center = np.array(data.shape) / 2
    center = np.asarray(center, dtype=np.float64)
    for _ in range(ndim):
        center = np.expand_dims(center, axis=-1)
    idx = np.indices(data.shape)
    r = np.sqrt(np.sum((idx - center) ** 2, axis=0))
    r = r.astype(int)
    r = np.where(r <= R, r, R + 1)
    tbin = np.bincount(r.ravel(), data.ravel())
    nr = np.bincount(r.ravel())
    with np.errstate(divide="ignore", invalid="ignore"):
        prof = tbin / nr
    return np.unique(r).ravel()[:-1], prof[:-1]


def export_structure(r, L, Nk=1001, radii=None):
    """utils.py:16-52 at eps -> 0."""
    L = np.asarray(L, dtype=np.float64)
    ndim = len(L)
    N = np.asarray(r).reshape(-1, ndim).shape[0]
    Sk = structure_factor(r, L, Nk, radii)
    gr = pair_correlation(Sk, L, N)
    center = int(Nk / 2)
    q, S = radial_profile(Sk, center=[center] * ndim)
    rq, g = radial_profile(gr, center=[center] * ndim)
    return q, S, Sk, rq / Nk, g, gr
