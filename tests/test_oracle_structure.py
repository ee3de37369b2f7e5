"""Pins the export_structure / radial_profile oracle (oracle/structure.py)
against brute force and closed forms (fresco/utils/utils.py:16-52,
138-152)."""
import math

import numpy as np
import pytest

from oracle import structure as ST


def brute_profile(data, center):
    data = np.asarray(data)
    R = int(data.shape[0] / 2)
    sums, counts = {}, {}
    for idx in np.ndindex(*data.shape):
        d2 = 0.0
        for a, i in enumerate(idx):
            d2 += (i - center[a]) ** 2
        b = int(math.floor(math.sqrt(d2)))
        b = b if b <= R else R + 1
        sums[b] = sums.get(b, 0.0) + float(data[idx])
        counts[b] = counts.get(b, 0) + 1
    bins = sorted(counts)
    return np.array(bins[:-1]), np.array([sums[b] / counts[b] for b in bins[:-1]])


@pytest.mark.parametrize("shape,center", [
    ((9,), [4]), ((10,), None), ((11, 11), [5, 5]), ((12, 12), None), ((7, 9), [3.3, 4.1]),
    ((6, 6, 6), [3, 3, 3]), ((5, 6, 7), None), ((8, 8, 8), [3.7, 4.2, 4.0]),
])
def test_radial_profile_vs_brute_force(shape, center):
    data = np.random.default_rng(len(shape)).normal(size=shape)
    q, p = ST.radial_profile(data, center=center)
    c = np.array(shape) / 2 if center is None else center
    qb, pb = brute_profile(data, c)
    assert np.array_equal(q, qb)
    assert np.allclose(p, pb, rtol=1e-12, atol=1e-14)


def test_radial_profile_constant():
    q, p = ST.radial_profile(np.full((21, 21), 2.5), center=[10, 10])
    assert np.array_equal(q, np.arange(11)) and np.allclose(p, 2.5)


def lattice(m, d):
    g = np.stack(np.meshgrid(*([np.arange(m) / m] * d), indexing="ij"), -1).reshape(-1, d)
    return g.ravel()


@pytest.mark.parametrize("d,m,Nk", [(2, 4, 8), (2, 3, 9), (2, 5, 11), (3, 2, 4), (3, 3, 7)])
def test_structure_factor_lattice(d, m, Nk):
    S = ST.structure_factor(lattice(m, d), np.ones(d), Nk)
    k = ST.mode_axis(Nk)
    on = np.ones([Nk] * d, dtype=bool)
    for a in range(d):
        sh = [1] * d
        sh[a] = Nk
        on &= (k.reshape(sh) % m) == 0
    assert np.allclose(S[on], m ** d, rtol=1e-12)
    assert np.max(np.abs(S[~on])) < 1e-10 * m ** d


def test_structure_factor_one_and_two_points_and_axis_order():
    assert np.allclose(ST.structure_factor(np.array([0.3, 0.7]), np.ones(2), 9), 1.0)
    dx = 0.1
    S = ST.structure_factor(np.array([0.2, 0.5, 0.2 + dx, 0.5]), np.ones(2), 9)
    k = ST.mode_axis(9)
    want = 1 + np.cos(2 * np.pi * dx * k)
    # the points differ along x: S varies with the first index only
    assert np.allclose(S, want[:, None] * np.ones((1, 9)), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("d,m,Nk", [(2, 4, 8), (2, 2, 6), (3, 2, 4)])
def test_pair_correlation_lattice(d, m, Nk):
    N = m ** d
    S = ST.structure_factor(lattice(m, d), np.ones(d), Nk)
    gr = ST.pair_correlation(S, np.ones(d), N)
    c = Nk // 2
    per = Nk // m
    for idx in np.ndindex(*gr.shape):
        r = np.array(idx) - c
        val = N * per ** d * float(np.all(r % per == 0)) - Nk ** d * float(np.all(r == 0))
        assert abs(gr[idx] - val / N) < 1e-9, (idx, gr[idx], val / N)


def test_weights_first_power_volumes():
    c = ST.weights(np.array([1.0, 2.0]), 2, 3)
    assert np.allclose(c, np.array([1.0, 8.0]) * 2 / 9)
