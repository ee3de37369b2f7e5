"""GPU export_structure / radial_profile (fresco/utils/utils.py:16-52,
138-152) against the oracle's exact sums and numpy restatement."""
import numpy as np
import pytest

from oracle import structure as ST

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def U():
    import fresco_amd.utils as U

    return U


def case(d, N, seed, radii=False):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0, 1, N * d)
    rad = rng.uniform(0.5, 1.5, N) if radii else None
    return r, rad


@pytest.mark.parametrize("d,N,Nk,radii", [(2, 300, 40, False), (2, 300, 33, True), (3, 200, 16, False),
                                          (3, 150, 15, True)])
def test_export_structure_vs_exact(U, d, N, Nk, radii):
    r, rad = case(d, N, seed=Nk, radii=radii)
    L = np.ones(d)
    print("export_structure d", d, "N", N, "Nk", Nk, flush=True)
    q, S, Sk, rq, g, gr = U.export_structure(r, L, Nk=Nk, radii=rad)
    q0, S0, Sk0, rq0, g0, gr0 = ST.export_structure(r, L, Nk=Nk, radii=rad)
    tol = 1e-9
    assert Sk.shape == (Nk,) * d and gr.shape == (Nk,) * d
    assert np.max(np.abs(Sk - Sk0)) <= 20 * tol * N
    assert np.array_equal(q, q0) and np.array_equal(rq, rq0)
    assert np.max(np.abs(S - S0)) <= 20 * tol * N
    # g(r) sums Nk^d values of S - 1 with weight V / N
    assert np.max(np.abs(gr - gr0)) <= 20 * tol * Nk ** d
    assert np.max(np.abs(g - g0)) <= 20 * tol * Nk ** d


def test_export_structure_fp32(U):
    r, _ = case(2, 500, seed=3)
    q, S, Sk, rq, g, gr = U.export_structure(r, np.ones(2), Nk=48, tol=1e-6, precision="fp32")
    Sk0 = ST.structure_factor(r, np.ones(2), 48)
    assert np.max(np.abs(Sk - Sk0)) <= 20 * 1e-6 * 500


@pytest.mark.parametrize("d,m,Nk", [(2, 4, 8), (2, 5, 20), (3, 2, 4), (3, 3, 9)])
def test_export_structure_lattice_closed_form(U, d, m, Nk):
    g = np.stack(np.meshgrid(*([np.arange(m) / m] * d), indexing="ij"), -1).reshape(-1, d)
    N = m ** d
    _, _, Sk, _, _, gr = U.export_structure(g.ravel(), np.ones(d), Nk=Nk)
    k = np.arange(Nk) - Nk // 2
    on = np.ones([Nk] * d, dtype=bool)
    for a in range(d):
        sh = [1] * d
        sh[a] = Nk
        on &= (k.reshape(sh) % m) == 0
    # NUFFT error: |d rho| <= ~eps * sum|c| = eps N, so |dS| <= ~2 eps N
    assert np.max(np.abs(Sk[on] - N)) <= 20 * 1e-9 * N and np.max(np.abs(Sk[~on])) < 1e-8 * N
    if Nk % m == 0:
        per = Nk // m
        c = Nk // 2
        idx = np.indices(gr.shape) - c
        want = (N * per ** d * np.all(idx % per == 0, axis=0) - Nk ** d * np.all(idx == 0, axis=0)) / N
        assert np.max(np.abs(gr - want)) < 1e-7 * Nk ** d


def test_axis_order_first_index_is_x(U):
    dx = 0.1
    _, _, Sk, _, _, _ = U.export_structure(np.array([0.2, 0.5, 0.2 + dx, 0.5]), np.ones(2), Nk=9)
    k = np.arange(9) - 4
    assert np.allclose(Sk, (1 + np.cos(2 * np.pi * dx * k))[:, None] * np.ones((1, 9)), atol=1e-8)


@pytest.mark.parametrize("shape,center", [((31,), None), ((40,), [20]), ((31, 31), None), ((40, 40), [20, 20]),
                                          ((17, 23), [8.3, 11.9]), ((17, 17, 17), None),
                                          ((20, 21, 22), [10.2, 9.7, 11.0])])
def test_radial_profile_vs_oracle(U, shape, center):
    data = np.random.default_rng(sum(shape)).normal(size=shape)
    q, p = U.radial_profile(data, center=center)
    q0, p0 = ST.radial_profile(data, center=center)
    assert np.array_equal(q, q0)
    assert np.allclose(p, p0, rtol=1e-12, atol=1e-14)
    q2, p2 = U.radial_profile(data, center=center)
    assert np.array_equal(p, p2)  # fixed summation order


def test_radial_profile_rejects_complex(U):
    with pytest.raises(TypeError):
        U.radial_profile(np.ones((4, 4)) + 1j)


def test_notebook_size(U):
    """The examples' call: N = 10000 2D points, Nk = K * 10 = 320 (even)."""
    import time

    r, _ = case(2, 10000, seed=11)
    t = time.perf_counter()
    U.export_structure(r, np.ones(2), Nk=320)  # first call: FFT kernels are built
    t1 = time.perf_counter()
    q, S, Sk, rq, g, gr = U.export_structure(r, np.ones(2), Nk=320)
    dt = time.perf_counter() - t1
    print("first call %.1f ms" % ((t1 - t) * 1e3), flush=True)
    q0, S0 = ST.radial_profile(Sk, center=[160, 160])
    rq0, g0 = ST.radial_profile(gr, center=[160, 160])
    print("export_structure N=10000 Nk=320: %.1f ms" % (dt * 1e3), flush=True)
    assert np.array_equal(q, q0) and np.allclose(S, S0, rtol=1e-12)
    assert np.array_equal(rq * 320, rq0) and np.allclose(g, g0, rtol=1e-10, atol=1e-10)
    # spot check of S at a few modes against the exact sum
    rr = r.reshape(-1, 2) * 2 * np.pi
    for (i1, i2) in [(160, 161), (170, 150), (0, 319), (300, 5)]:
        k = np.array([i1 - 160, i2 - 160])
        rho = np.sum(np.exp(1j * rr @ k))
        assert abs(Sk[i1, i2] - abs(rho) ** 2 / 10000) < 20 * 1e-9 * 10000
