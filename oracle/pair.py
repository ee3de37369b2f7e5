"""Brute-force restatement of the reference's pair potentials.

TEST INFRASTRUCTURE ONLY (the product path never imports this module).

  InversePowerInteraction  source/fresco/inversepower_potential.hpp:18-78
  WCA_interaction          source/fresco/wca.hpp:14-82
  all-pairs accumulation   source/fresco/simple_pairwise_potential.hpp:57-130
  minimum image            source/fresco/distance.hpp:32-36 (round_fast =
                           round half to even), 172-192
"""
from __future__ import annotations

import numpy as np


def _pairs(x, d, box):
    x = np.asarray(x, dtype=np.float64).reshape(-1, d)
    n = len(x)
    i, j = np.tril_indices(n, -1)  # atom_j < atom_i (simple_pairwise_potential.hpp:60-62)
    dr = x[i] - x[j]
    if box is not None:
        box = np.asarray(box, dtype=np.float64)
        dr = dr - np.rint(dr * (1.0 / box)) * box
    return i, j, dr


def inverse_power_eg(r2, rs, a, eps):
    """(E, gij) per pair, gij = -(dV/dr)/r (inversepower_potential.hpp:44-58)."""
    inside = r2 < rs * rs
    r = np.sqrt(np.where(inside, r2, 0.0))
    f = np.where(inside, np.power(np.where(inside, 1 - r / rs, 1.0), a) * eps, 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = np.where(inside, -f / ((r - rs) * r), 0.0)
    return f / a, g


def wca_eg(r2, sig, eps):
    """(E, gij) per pair (wca.hpp:34-63)."""
    C6 = sig ** 6
    C12 = C6 * C6
    coff2 = 2.0 ** (1.0 / 3) * sig * sig
    ir2 = 1.0 / r2
    ir6 = ir2 ** 3
    ir12 = ir6 * ir6
    inside = r2 < coff2
    E = np.where(inside, 4 * eps * (-C6 * ir6 + C12 * ir12) + eps, 0.0)
    g = np.where(inside, 4 * eps * (-6 * C6 * ir6 + 12 * C12 * ir12) * ir2, 0.0)
    return E, g


def energy_gradient(kind, p1, eps, radii, x, d, box=None):
    """Total energy and gradient (grad[i] -= gij dr, grad[j] += gij dr)."""
    radii = np.asarray(radii, dtype=np.float64)
    i, j, dr = _pairs(x, d, box)
    r2 = np.sum(dr * dr, axis=1)
    if kind == "inverse_power":
        E, g = inverse_power_eg(r2, radii[i] + radii[j], p1, eps)
    else:
        E, g = wca_eg(r2, p1, eps)
    n = len(radii)
    grad = np.zeros((n, d))
    f = g[:, None] * dr
    np.add.at(grad, i, -f)
    np.add.at(grad, j, f)
    return float(np.sum(E)), grad.ravel()
