"""The vector-free LBFGS recursion (the two-loop recursion of
lbfgs.cpp:128-160 run on the coefficients of the memory vectors and g, Chen,
Wang & Zhou 2014) against the slot-by-slot form: the same iterates up to
rounding, recorded-graph replays bitwise equal to launches, backtracking,
every instantiated memory width, and set_func_gradient."""
import ctypes

import numpy as np
import pytest

from oracle import sf_direct as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from fresco_amd.optimize import LBFGS_CPP
    import fresco_amd.potentials as P

    return LBFGS_CPP, P


def stealthy(kp, d, kmax):
    k = D.mode_grid(kp, d)
    kk = np.linalg.norm(k, axis=1)
    V = ((kk > 0) & (kk <= kmax)).astype(float)
    return np.zeros_like(V), V


def nuwu(P, N, d, kp, kmax, prec, eps, seed=0):
    Sk, V = stealthy(kp, d, kmax)
    pot = P.NUwU(np.ones(N), kp, Sk, V, np.ones(d), eps, precision=prec)
    return pot, np.random.default_rng(seed).uniform(0, 1, d * N)


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def test_vector_free_tracks_two_loop_3d(mods):
    """3D NUwU, N = 30000 (90000 unknowns: the automatic choice is the
    vector-free form), M = 10, 40 iterations."""
    LBFGS_CPP, P = mods
    pot, x0 = nuwu(P, 30000, 3, 20, 8, "fp64-adjoint", 1e-10)
    a = LBFGS_CPP(x0, pot, tol=0.0, M=10)  # auto
    b = LBFGS_CPP(x0, pot, tol=0.0, M=10, recursion="two-loop")
    E0 = None
    for it in range(40):
        ra = a.one_iteration()
        rb = b.one_iteration()
        E0 = E0 or ra.energy
        assert ra.nfev == rb.nfev
        assert rel(ra.coords - x0, rb.coords - x0) < 1e-8, it
        assert abs(ra.energy - rb.energy) <= 1e-10 * E0, it
        if ra.energy > 1e-6 * E0:
            # near the minimum H0 = ys / yy is a ratio of rounding-level terms
            assert abs(ra["H0"] - rb["H0"]) <= 1e-8 * abs(rb["H0"]), it
    print("E", E0, "->", ra.energy, "H0", ra["H0"], rb["H0"], flush=True)
    assert a.graph_stats() == (1, 40)


@pytest.mark.parametrize("M", [1, 3, 7, 12, 16])
def test_vector_free_widths(mods, M):
    LBFGS_CPP, P = mods
    pot, x0 = nuwu(P, 600, 2, 12, 6, "fp64", 1e-12, seed=M)
    a = LBFGS_CPP(x0, pot, tol=0.0, M=M, recursion="vector-free")
    b = LBFGS_CPP(x0, pot, tol=0.0, M=M, recursion="two-loop")
    for it in range(3 * M + 5):
        ra = a.one_iteration()
        rb = b.one_iteration()
        assert ra.nfev == rb.nfev
        assert rel(ra.coords - x0, rb.coords - x0) < 1e-9, (M, it)


def test_vector_free_graph_bitwise_with_backtracking(mods):
    LBFGS_CPP, P = mods
    pot, x0 = nuwu(P, 400, 2, 12, 6, "fp64", 1e-12, seed=7)
    kw = dict(tol=0.0, M=3, maxstep=1.0, maxErise=-0.6, rel_energy=True, recursion="vector-free")
    ga = LBFGS_CPP(x0, pot, use_graph=True, **kw)
    eb = LBFGS_CPP(x0, pot, use_graph=False, **kw)
    xr = np.random.default_rng(1).uniform(0, 1, len(x0))
    for it in range(40):
        ra = ga.one_iteration()
        pot.get_energy_gradient(xr)
        rb = eb.one_iteration()
        assert np.array_equal(ra.coords, rb.coords) and ra.energy == rb.energy
        assert ra["H0"] == rb["H0"] and ra.nfev == rb.nfev
    assert ra.nfev > ra.nsteps + 20
    assert ga.graph_stats() == (1, 40)


def test_vector_free_set_func_gradient(mods):
    """Replacing f, g by the same values mid-run recomputes the memory's
    products with g; the run continues exactly as without the call."""
    from fresco_amd import _lib

    LBFGS_CPP, P = mods
    pot, x0 = nuwu(P, 500, 2, 12, 6, "fp64", 1e-12, seed=3)
    a = LBFGS_CPP(x0, pot, tol=0.0, M=4, recursion="vector-free")
    b = LBFGS_CPP(x0, pot, tol=0.0, M=4, recursion="vector-free")
    for _ in range(6):
        ra = a.one_iteration()
        rb = b.one_iteration()
    g = np.ascontiguousarray(rb.grad, dtype=np.float64)
    _lib.check(_lib.load().fresco_opt_set_func_gradient(b._h, float(rb.energy), _lib.dptr(g), len(g)))
    for _ in range(10):
        ra = a.one_iteration()
        rb = b.one_iteration()
        assert np.array_equal(ra.coords, rb.coords) and ra.energy == rb.energy


def test_form_rules(mods):
    LBFGS_CPP, P = mods
    pot, x0 = nuwu(P, 100, 2, 8, 4, "fp64", 1e-10)
    with pytest.raises(ValueError):
        LBFGS_CPP(x0, pot, M=17, recursion="vector-free")
    with pytest.raises(ValueError):
        LBFGS_CPP(x0, pot, recursion="sideways")
    opt = LBFGS_CPP(x0, pot, M=4)
    opt.one_iteration()
    with pytest.raises(ValueError):
        opt._set(__import__("fresco_amd")._lib.OPT_LBFGS_FORM, 2)
