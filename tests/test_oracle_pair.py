"""The pair-potential oracle against the reference's known answers
(cpp_tests/source/test_inversepower.cpp:17-46), central differences, and the
reference's own headers compiled from source (oracle/_ref)."""
import numpy as np
import pytest

from oracle import pair as OP
from oracle import ref_pair as RP

# test_inversepower.cpp:20-44
X_IP = np.array([0.1, 0.2, 0.3, 0.44, 0.55, 1.66, 0.88, 1.1, 3.32])
RADII = np.array([0.91, 1.1, 1.13])
ETRUE_CART = 0.023173380132354017
ETRUE_PER = 0.60660574285698954


def fd(f, x, h=1e-6):
    g = np.zeros_like(x)
    for k in range(x.size):
        e = np.zeros_like(x)
        e[k] = h
        g[k] = (f(x + e) - f(x - e)) / (2 * h)
    return g


def test_inverse_power_known_answers():
    E, g = OP.energy_gradient("inverse_power", 2.5, 1.0, RADII, X_IP, 3)
    assert E == pytest.approx(ETRUE_CART, abs=1e-10)
    Ep, gp = OP.energy_gradient("inverse_power", 2.5, 1.0, RADII, X_IP, 3, box=np.ones(3))
    assert Ep == pytest.approx(ETRUE_PER, abs=1e-10)
    for box, gg in ((None, g), (np.ones(3), gp)):
        num = fd(lambda y: OP.energy_gradient("inverse_power", 2.5, 1.0, RADII, y, 3, box=box)[0], X_IP)
        assert np.allclose(gg, num, atol=1e-6)


def test_wca_fd_and_cutoff():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 3, 30)
    box = np.full(3, 3.0)
    E, g = OP.energy_gradient("wca", 1.0, 0.7, np.ones(10), x, 3, box=box)
    num = fd(lambda y: OP.energy_gradient("wca", 1.0, 0.7, np.ones(10), y, 3, box=box)[0], x)
    assert np.allclose(g, num, rtol=1e-5, atol=1e-6 * np.abs(g).max())
    # one pair at the cutoff 2^(1/6) sigma: energy 0 from both sides (continuous)
    rc = 2 ** (1 / 6)
    e_in = OP.energy_gradient("wca", 1.0, 1.0, np.ones(2), np.array([0, 0, 0, rc - 1e-9, 0, 0.0]), 3)[0]
    assert abs(e_in) < 1e-7


@pytest.mark.skipif(not RP.available(), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("kind,p1", [("inverse_power", 2.5), ("wca", 0.9)])
@pytest.mark.parametrize("d", [2, 3])
def test_oracle_vs_reference_build(kind, p1, d):
    rng = np.random.default_rng(10 + d)
    n = 60
    radii = rng.uniform(0.3, 0.5, n)
    box = np.full(d, 4.0)
    x = rng.uniform(-2, 2, n * d)
    for method, b in (("cartesian", None), ("periodic", box)):
        E, g = OP.energy_gradient(kind, p1, 1.3, radii, x, d, box=b)
        R = RP.RefPair(kind, method, p1, 1.3, radii, box)
        Er, gr = R.get_energy_gradient(x)
        assert E == pytest.approx(Er, rel=1e-12, abs=1e-14)
        assert np.allclose(g, gr, rtol=1e-11, atol=1e-13)
    if kind == "inverse_power":
        # the reference's cell lists agree with its all-pairs sum (rcut = 2 max r)
        for scale in (1.0, 2.0):
            Rc = RP.RefPair(kind, "cells", p1, 1.3, radii, box, scale)
            Ec, gc = Rc.get_energy_gradient(x)
            E, g = OP.energy_gradient(kind, p1, 1.3, radii, x, d, box=box)
            assert Ec == pytest.approx(E, rel=1e-12)
            assert np.allclose(gc, g, rtol=1e-11, atol=1e-13)


@pytest.mark.skipif(not RP.available(), reason="oracle/_ref not built (needs the reference sources)")
def test_reference_build_known_answers():
    assert RP.RefPair("inverse_power", "cartesian", 2.5, 1.0, RADII, np.ones(3)).get_energy(X_IP) == pytest.approx(
        ETRUE_CART, abs=1e-10)
    assert RP.RefPair("inverse_power", "periodic", 2.5, 1.0, RADII, np.ones(3)).get_energy(X_IP) == pytest.approx(
        ETRUE_PER, abs=1e-10)
