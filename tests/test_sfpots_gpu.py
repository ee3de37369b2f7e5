"""GPU parity of UwU, UwNU and NUwNU (libfresco_hip through the C ABI) against
the oracle's exact restatements.

Tolerances (relative l2 unless stated):
  UwU   fp64 1e-12, fp32 1e-5                      (FFT only, no kernel error)
  UwNU  fp64 at eps=1e-12: 1e-9; fp32 at eps=1e-6: 1e-5
  NUwNU fp64 1e-11 (exact sums), fp32 1e-5
"""
import math

import numpy as np
import pytest

from oracle import sf_direct as O

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def P():
    import fresco_amd.potentials as P

    return P


# ---------------------------------------------------------------------------
# UwU
# ---------------------------------------------------------------------------

UWU_TOL = {"fp64": 1e-12, "fp32": 1e-5}


def uwu_case(shape, seed):
    rng = np.random.default_rng(seed)
    Sk = rng.uniform(0.5, 3, shape)
    Sk.flat[::4] = 0.0
    V = rng.uniform(0, 1, shape)
    V.flat[::7] = 0.0
    x = rng.uniform(0.2, 1.8, int(np.prod(shape)))
    return Sk, V, x


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(64,), (12, 10), (9, 8, 7), (32, 32, 32)])
def test_uwu_parity(P, prec, shape):
    Sk, V, x = uwu_case(shape, seed=sum(shape))
    L = np.ones(len(shape))
    ref = O.UwU(np.array(shape), 0, Sk, V, L)
    E0, g0 = ref.get_energy_gradient(x)
    pot = P.UwU(np.array(shape), 0, Sk, V, L, precision=prec)
    E, g = pot.get_energy_gradient(x)
    print("uwu", shape, prec, E, E0, flush=True)
    assert abs(E - E0) / abs(E0) < UWU_TOL[prec]
    assert rel(g, g0) < UWU_TOL[prec]
    assert abs(pot.get_energy(x) - E0) / abs(E0) < UWU_TOL[prec]
    rho = O.dft_grid(x.reshape(shape), +1).ravel()
    assert rel(pot.get_rho(), rho) < UWU_TOL[prec]


def test_uwu_exact_gradient_matches_finite_differences(P):
    shape = (6, 5, 4)
    Sk, V, x = uwu_case(shape, seed=3)
    pot = P.UwU(np.array(shape), 0, Sk, V, np.ones(3), exact_grad=True)
    _, g = pot.get_energy_gradient(x)
    h = 1e-6
    fd = np.empty_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e[i] = h
        fd[i] = (pot.get_energy(x + e) - pot.get_energy(x - e)) / (2 * h)
    assert rel(g, fd) < 1e-6
    # and the relation to the reference gradient: exact = 2 ref - const
    _, g_ref = P.UwU(np.array(shape), 0, Sk, V, np.ones(3)).get_energy_gradient(x)
    c = 2 * g_ref - g
    assert np.ptp(c) < 1e-10 * np.abs(g).max()


# ---------------------------------------------------------------------------
# UwNU
# ---------------------------------------------------------------------------

NU_TOL = {"fp64": 1e-9, "fp32": 1e-5}
NU_EPS = {"fp64": 1e-12, "fp32": 1e-6}


def uwnu_case(shape, nk, seed, kmax):
    rng = np.random.default_rng(seed)
    d = len(shape)
    K = rng.uniform(-kmax, kmax, (nk, d))
    Sk = rng.uniform(0.5, 3, nk)
    Sk[::5] = 0.0
    V = rng.uniform(0, 1, nk)
    x = rng.uniform(0.2, 1.8, int(np.prod(shape)))
    return K, Sk, V, x


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape,nk,L", [((40,), 70, (1.0,)), ((12, 9), 150, (1.0, 2.0)), ((8, 7, 6), 200, (1.0, 1.0, 0.5))])
def test_uwnu_parity(P, prec, shape, nk, L):
    # K up to 1.3 L: k = 2 pi K / L spans several periods of the mode lattice
    K, Sk, V, x = uwnu_case(shape, nk, seed=nk, kmax=1.3)
    K = K * np.asarray(L)
    L = np.asarray(L)
    ref = O.UwNU(np.array(shape), K, Sk, V, L)
    E0, g0 = ref.get_energy_gradient(x)
    pot = P.UwNU(np.array(shape), K, Sk, V, L, NU_EPS[prec], precision=prec)
    E, g = pot.get_energy_gradient(x)
    print("uwnu", shape, prec, pot.plan_info(), E, E0, flush=True)
    assert abs(E - E0) / abs(E0) < NU_TOL[prec]
    assert rel(g, g0) < NU_TOL[prec]
    assert rel(pot.get_rho(), ref.rho(x)) < NU_TOL[prec]
    assert abs(pot.get_energy(x) - E0) / abs(E0) < NU_TOL[prec]


def test_uwnu_integer_k_is_a_dft(P):
    # integer K on a grid of the same size: rho is the plain DFT of the grid
    shape = (8, 6)
    rng = np.random.default_rng(4)
    kx, ky = np.meshgrid(np.arange(-4, 4), np.arange(-3, 3), indexing="ij")
    K = np.stack([kx.ravel(), ky.ravel()], 1).astype(float)
    nk = len(K)
    x = rng.uniform(0.5, 1.5, 48)
    pot = P.UwNU(np.array(shape), K, np.ones(nk), np.ones(nk), np.ones(2), 1e-12)
    pot.get_energy(x)
    m = O.uwnu_mode_index(shape)
    rho = np.exp(1j * TWO_PI * K @ m.T.astype(float)) @ x
    assert rel(pot.get_rho(), rho) < 1e-10


# ---------------------------------------------------------------------------
# NUwNU
# ---------------------------------------------------------------------------

# (method, precision): (eps, tolerance); "direct" sums are exact up to rounding
NUNU_CASES = {("direct", "fp64"): (1e-6, 1e-11), ("direct", "fp32"): (1e-6, 1e-5),
              ("nufft", "fp64"): (1e-12, 1e-9), ("nufft", "fp32"): (1e-6, 1e-5)}


@pytest.mark.parametrize("method,prec", sorted(NUNU_CASES))
@pytest.mark.parametrize("d,N,nk", [(1, 300, 90), (2, 500, 300), (3, 700, 400)])
def test_nuwnu_parity(P, method, prec, d, N, nk):
    rng = np.random.default_rng(100 + d)
    radii = rng.uniform(0.5, 1.5, N)
    K = rng.uniform(-20, 20, (nk, d))
    Sk = rng.uniform(0, 2, nk)
    Sk[::3] = 0.0
    V = rng.uniform(0, 1, nk)
    L = np.linspace(1.0, 1.5, d)
    p = rng.uniform(-0.4, 1.4, N * d)
    ref = O.NUwNU(radii, K.ravel(), Sk, V, L)
    E0, g0 = ref.get_energy_gradient(p)
    eps, tol = NUNU_CASES[(method, prec)]
    pot = P.NUwNU(radii, K.ravel(), Sk, V, L, eps, precision=prec, method=method)
    E, g = pot.get_energy_gradient(p)
    print("nuwnu", method, d, N, nk, prec, pot.plan_info(), E, E0, flush=True)
    assert abs(E - E0) / abs(E0) < tol
    assert rel(g, g0) < tol
    assert rel(pot.get_rho(), ref.rho(p)) < tol
    assert abs(pot.get_energy(p) - E0) / abs(E0) < tol


def test_nuwnu_mode_count_follows_k_and_sizes_checked(P):
    K = np.arange(12, dtype=float).reshape(6, 2)
    pot = P.NUwNU(np.ones(5), K.ravel(), np.ones(8), np.ones(8), np.ones(2))
    assert pot.nmodes == 6
    with pytest.raises(ValueError):
        P.NUwNU(np.ones(5), K.ravel(), np.ones(4), np.ones(4), np.ones(2))
    with pytest.raises(ValueError):
        P.UwNU(np.array([4, 4]), K.ravel(), np.ones(7), np.ones(7), np.ones(2))
    with pytest.raises(ValueError):
        P.UwU(np.array([4, 4]), 0, np.ones(15), np.ones(15), np.ones(2))


def test_nuwnu_type3_offcentre_k_and_fd(P):
    """k box far from the origin (centre D != 0) and unequal box lengths:
    the type-3 centring phases; fp64 gradient against central differences."""
    rng = np.random.default_rng(77)
    d, N, nk = 2, 400, 250
    K = rng.uniform(30, 55, (nk, d)) * np.array([1.0, -1.0])
    Sk = rng.uniform(0.5, 2, nk)
    V = rng.uniform(0, 1, nk)
    L = np.array([1.0, 1.3])
    p = rng.uniform(0, 1.3, N * d)
    radii = np.ones(N)
    ref = O.NUwNU(radii, K.ravel(), Sk, V, L)
    E0, g0 = ref.get_energy_gradient(p)
    pot = P.NUwNU(radii, K.ravel(), Sk, V, L, 1e-12)
    E, g = pot.get_energy_gradient(p)
    assert abs(E - E0) / abs(E0) < 1e-9
    assert rel(g, g0) < 1e-9
    v = rng.normal(size=p.size)
    h = 1e-6
    fd = (pot.get_energy(p + h * v) - pot.get_energy(p - h * v)) / (2 * h)
    gp = g * np.tile(2 * math.pi / L, N)
    assert abs(fd - np.dot(gp, v)) / abs(fd) < 1e-6
