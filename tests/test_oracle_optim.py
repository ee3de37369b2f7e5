"""Pins the optimizer oracle (oracle/optim.py) against the reference's own
lbfgs.cpp / modified_fire.cpp compiled from its sources (oracle/_ref), and
against the analytic minimum of a harmonic potential (CPU only)."""
import numpy as np
import pytest

from oracle import optim as OP
from oracle import ref_opt as RO
from oracle import sf_direct as D

needs_ref = pytest.mark.skipif(not RO.available(), reason="oracle/_ref not built (reference sources absent)")


class Harmonic:
    """E = 1/2 |x|^2 (cpp_tests/source/test_base_potential.cpp:11-25)."""

    def get_energy_gradient(self, x):
        x = np.asarray(x)
        return 0.5 * float(np.dot(x, x)), x.copy()


def small_nuwu(seed=0):
    rng = np.random.default_rng(seed)
    N, d, n = 12, 2, 9
    M = n * n
    V = rng.uniform(0, 1, M)
    V[M // 2] = 0
    Sk = np.zeros(M)
    pot = D.NUwU(np.ones(N), 0, Sk, V, np.ones(d))
    return pot, rng.uniform(0, 1, N * d)


def test_lbfgs_harmonic_minimum():
    o = OP.LBFGS(np.array([1.0, 1.5]), Harmonic(), tol=1e-10)
    o.run(100)
    assert np.allclose(o.x, 0, atol=1e-9)
    assert o.rms <= 1e-10


@needs_ref
@pytest.mark.parametrize("M,rel", [(4, False), (10, False), (4, True)])
def test_lbfgs_matches_reference_binary(M, rel):
    pot, x0 = small_nuwu(M)
    a = OP.LBFGS(x0, pot, tol=1e-12, M=M)
    a.use_relative_f = rel
    b = RO.LBFGS(x0, pot, tol=1e-12, M=M, rel=rel)
    a.initialize_func_gradient()
    for it in range(40):
        a.one_iteration()
        b.one_iteration()
        s = b.state()
        assert s["nfev"] == a.nfev and s["niter"] == a.iter
        assert np.allclose(s["x"], a.x, rtol=1e-9, atol=1e-12), it
        assert np.isclose(s["f"], a.f, rtol=1e-9)
        assert np.isclose(s["H0"], a.H0, rtol=1e-7)


@needs_ref
def test_fire_matches_reference_binary():
    pot, x0 = small_nuwu(3)
    a = OP.ModifiedFire(x0, pot, dtstart=0.01, dtmax=0.1, maxstep=0.05, tol=1e-12)
    b = RO.ModifiedFire(x0, pot, dtstart=0.01, dtmax=0.1, maxstep=0.05, tol=1e-12)
    a.initialize_func_gradient()
    b.stop_criterion_satisfied()
    for it in range(60):
        a.one_iteration()
        b.one_iteration()
        s = b.state()
        assert s["nfev"] == a.nfev and s["niter"] == a.iter
        assert np.allclose(s["x"], a.x, rtol=1e-9, atol=1e-12), it
        assert np.isclose(s["f"], a.f, rtol=1e-9)


@needs_ref
def test_reference_run_converges_harmonic():
    b = RO.LBFGS(np.array([1.0, 1.5]), Harmonic(), tol=1e-10)
    b.run(100)
    assert np.allclose(b.state()["x"], 0, atol=1e-9)
