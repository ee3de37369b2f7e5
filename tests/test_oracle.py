"""Pins the CPU oracle (oracle/sf_direct.py) against closed forms, brute force
and finite differences, so that a dropped term, a wrong sign or index, or a
transposed operand in the oracle fails here (CPU only)."""
import math

import numpy as np
import pytest

from oracle import sf_direct as O

TWO_PI = 2 * math.pi


def fd_grad(f, p, h=1e-6):
    g = np.zeros_like(p)
    for i in range(len(p)):
        pp = p.copy()
        pm = p.copy()
        pp[i] += h
        pm[i] -= h
        g[i] = (f(pp) - f(pm)) / (2 * h)
    return g


# --------------------------------------------------------------------------
# conventions
# --------------------------------------------------------------------------

def test_c_round_half_away_from_zero():
    assert list(O.c_round([0.5, 1.5, 2.5, -0.5, -1.5, 0.49])) == [1, 2, 3, -1, -2, 0]


@pytest.mark.parametrize(
    "m,d,kp",
    [(129, 1, 64), (128, 1, 63), (65 ** 2, 2, 32), (64 ** 2, 2, 31), (129 ** 3, 3, 64), (128 ** 3, 3, 63),
     (17 ** 3, 3, 8), (1, 3, 0)],
)
def test_kvec_size_closed_form(m, d, kp):
    # nuwu.hpp:54-79: even grids silently truncate to the next odd one below
    assert O.kvec_size(m, d) == kp


def test_mode_grid_order():
    k = O.mode_grid(1, 3)
    # x fastest, then y, then z (nuwu.hpp:277,284-285)
    assert k[0].tolist() == [-1, -1, -1]
    assert k[1].tolist() == [0, -1, -1]
    assert k[3].tolist() == [-1, 0, -1]
    assert k[9].tolist() == [-1, -1, 0]
    assert k[13].tolist() == [0, 0, 0]
    k2 = O.mode_grid(2, 2)
    assert k2[5].tolist() == [-2, -1]


def test_weights_hand_computed():
    r = np.array([1.0, 2.0])
    c2 = O.nuwu_weights(r, 2)  # (pi r^2)^2 -> pi^2 {1, 16}, normalised to sum 2
    assert np.allclose(c2, 2 * np.array([1, 16]) / 17)
    c3 = O.nuwu_weights(r, 3)  # ((4/3) pi r^3)^3 -> {1, 512}
    assert np.allclose(c3, 2 * np.array([1, 512]) / 513)
    assert np.allclose(O.nuwu_weights(r, 1), [1, 1])


def test_coordinate_maps():
    p = np.array([0.25, 0.75, -0.6, 1.5])
    L = np.ones(1)
    xc = O.coords_centered(p, L).ravel()
    xs = O.coords_shifted(p, L).ravel()
    assert np.allclose(xc, TWO_PI * np.array([0.25, -0.25, 0.4, -0.5]))
    assert np.allclose(xs, xc + TWO_PI)


# --------------------------------------------------------------------------
# non-uniform sums
# --------------------------------------------------------------------------

@pytest.mark.parametrize("d", [1, 2, 3])
def test_separable_equals_dense(d):
    rng = np.random.default_rng(1)
    x = rng.uniform(-4, 4, (23, d))
    c = rng.normal(size=23) + 1j * rng.normal(size=23)
    kp = 3
    k = O.mode_grid(kp, d)
    for sign in (+1, -1):
        assert np.allclose(O.type1_grid(x, c, kp, sign), O.nudft_dense(x, c, k, sign), atol=1e-12)
        f = rng.normal(size=len(k)) + 1j * rng.normal(size=len(k))
        assert np.allclose(O.type2_grid(x, f, kp, sign), O.nudft_dense_adj(x, f, k, sign), atol=1e-12)


def test_single_point_and_pair():
    x = np.array([[0.3, 1.1]])
    rho = O.type1_grid(x, np.array([2.5]), 4, +1)
    assert np.allclose(np.abs(rho), 2.5)
    x2 = np.array([[0.3, 1.1], [2.0, -0.4]])
    rho2 = O.type1_grid(x2, np.ones(2), 4, +1)
    k = O.mode_grid(4, 2)
    dx = x2[1] - x2[0]
    assert np.allclose(np.abs(rho2) ** 2, 2 + 2 * np.cos(k @ dx))


def test_square_lattice_bragg_peaks():
    # n x n lattice in the unit box: rho_k = N on k = 0 mod n, else 0
    n = 4
    g = (np.arange(n) + 0.5) / n
    p = np.stack(np.meshgrid(g, g, indexing="ij"), axis=-1).reshape(-1, 2)
    x = O.coords_centered(p.ravel(), np.ones(2))
    kp = 6
    rho = O.type1_grid(x, np.ones(n * n), kp, +1)
    k = O.mode_grid(kp, 2)
    on = np.all(k % n == 0, axis=1)
    assert np.allclose(np.abs(rho[on]), n * n)
    assert np.allclose(np.abs(rho[~on]), 0, atol=1e-10)


def test_hermitian_and_adjoint():
    rng = np.random.default_rng(2)
    x = rng.uniform(0, TWO_PI, (31, 3))
    c = rng.uniform(0.5, 2, 31)
    kp = 2
    rho = O.type1_grid(x, c, kp, +1)
    assert np.allclose(rho[::-1], np.conj(rho))  # rho_{-k} = conj(rho_k)
    f = rng.normal(size=len(rho)) + 1j * rng.normal(size=len(rho))
    cz = rng.normal(size=31) + 1j * rng.normal(size=31)
    # <f, T1(+1) c> = <T2(-1) f, c>
    lhs = np.vdot(f, O.type1_grid(x, cz, kp, +1))
    rhs = np.vdot(O.type2_grid(x, f, kp, -1), cz)
    assert np.isclose(lhs, rhs)


# --------------------------------------------------------------------------
# energies: closed forms
# --------------------------------------------------------------------------

def test_nuwu_lattice_energy_closed_form():
    n = 4
    g = (np.arange(n) + 0.5) / n
    p = np.stack(np.meshgrid(g, g, indexing="ij"), axis=-1).ravel()
    N = n * n
    kp = 6
    M = (2 * kp + 1) ** 2
    V = np.ones(M)
    # S0 = 0: E = sum over reciprocal-lattice modes of (N^2/N)^2 = count * N^2
    pot = O.NUwU(np.ones(N), 0, np.zeros(M), V, np.ones(2))
    count = len([1 for a in range(-kp, kp + 1) for b in range(-kp, kp + 1) if a % n == 0 and b % n == 0])
    assert np.isclose(pot.get_energy(p), count * N ** 2)
    # S0 = N on the peaks, 1 elsewhere: peaks contribute 0, others ((0-1)/1)^2 = 1
    k = O.mode_grid(kp, 2)
    on = np.all(k % n == 0, axis=1)
    Sk = np.where(on, N, 1.0)
    pot2 = O.NUwU(np.ones(N), 0, Sk, V, np.ones(2))
    assert np.isclose(pot2.get_energy(p), M - count)


# --------------------------------------------------------------------------
# gradients: finite differences (reference convention: d/dx, x = 2 pi p / L)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("d", [1, 2, 3])
def test_nuwu_gradient_fd(d):
    rng = np.random.default_rng(3 + d)
    N = 9
    n = 7 if d == 3 else 9
    M = n ** d
    radii = rng.uniform(0.6, 1.4, N)
    Sk = rng.uniform(0, 2, M)
    Sk[::3] = 0
    V = rng.uniform(0, 1, M)
    L = np.full(d, 1.0)
    pot = O.NUwU(radii, 0, Sk, V, L)
    p = rng.uniform(0, 1, N * d)
    E, g = pot.get_energy_gradient(p)
    assert np.isclose(E, pot.get_energy(p))
    fd = fd_grad(pot.get_energy, p) * L[0] / TWO_PI
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-7 * np.abs(g).max())


def test_nuwu_gradient_sums_to_zero_uniform_c():
    rng = np.random.default_rng(7)
    pot = O.NUwU(np.ones(12), 0, np.zeros(81), rng.uniform(0, 1, 81), np.ones(2))
    _, g = pot.get_energy_gradient(rng.uniform(0, 1, 24))
    assert np.allclose(g.reshape(-1, 2).sum(0), 0, atol=1e-9 * np.abs(g).max())


def test_nuwnu_gradient_fd_and_missing_conj_c():
    rng = np.random.default_rng(8)
    N, nk, d = 8, 13, 2
    K = rng.uniform(-6, 6, (nk, d))
    Sk = rng.uniform(0, 3, nk)
    Sk[::4] = 0
    V = rng.uniform(0, 1, nk)
    p = rng.uniform(0, 1, N * d)
    # uniform radii: exact derivative
    pot = O.NUwNU(np.ones(N), K.ravel(), Sk, V, np.ones(d))
    E, g = pot.get_energy_gradient(p)
    fd = fd_grad(pot.get_energy, p) / TWO_PI
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-8 * np.abs(g).max())
    # non-uniform radii: the reference omits conj(c_j) (nuwnu.hpp:269): g = exact / c
    radii = rng.uniform(0.7, 1.3, N)
    pot2 = O.NUwNU(radii, K.ravel(), Sk, V, np.ones(d))
    _, g2 = pot2.get_energy_gradient(p)
    fd2 = fd_grad(pot2.get_energy, p) / TWO_PI
    c = O.nuwu_weights(radii, d)
    assert np.allclose(g2 * np.repeat(c, d), fd2, rtol=1e-6, atol=1e-8 * np.abs(fd2).max())


def test_uwu_gradient_quirk_relation():
    # reference grad = Re(FFT_-1((2/sumI) V D rho / S0)); the exact derivative is
    # 2*grad - C with the constant C = sum_k 2 V D |rho|^2 / (S0 sumI^2) from d(sumI)
    rng = np.random.default_rng(9)
    shape = (5, 6)
    Sk = rng.uniform(0.5, 3, shape)
    Sk[0, 1] = 0
    V = rng.uniform(0, 1, shape)
    pot = O.UwU(np.array(shape), 0, Sk, V, np.ones(2))
    x = rng.uniform(0.5, 1.5, 30)
    E, g = pot.get_energy_gradient(x)
    fd = fd_grad(pot.get_energy, x)
    rho = O.dft_grid(x.reshape(shape), +1).ravel()
    sI = x.sum()
    S0 = np.fft.ifftshift(Sk).ravel()
    Vv = np.fft.ifftshift(V).ravel()
    D = np.abs(rho) ** 2 / sI - S0
    nz = S0 != 0
    D[nz] /= S0[nz]
    q = 2 * Vv * D * np.abs(rho) ** 2 / sI ** 2
    q[nz] /= S0[nz]
    assert np.allclose(fd, 2 * g - q.sum(), rtol=1e-6, atol=1e-8 * np.abs(fd).max())
    # brute-force DFT pins the FFTW +1 convention of dft_grid
    m = np.stack(np.meshgrid(np.arange(5), np.arange(6), indexing="ij"), -1).reshape(-1, 2)
    ph = 2 * np.pi * (m[:, None, 0] * m[None, :, 0] / 5 + m[:, None, 1] * m[None, :, 1] / 6)
    assert np.allclose(np.exp(1j * ph) @ x, rho)


def test_uwnu_gradient_is_fd_over_two_pi():
    rng = np.random.default_rng(10)
    shape = (5, 4)
    nk, d = 6, 2
    K = rng.uniform(-0.4, 0.4, (nk, d))
    Sk = rng.uniform(1, 3, nk)
    Sk[2] = 0
    V = rng.uniform(0, 1, nk)
    pot = O.UwNU(np.array(shape), K, Sk, V, np.ones(d))
    x = rng.uniform(0.5, 1.5, 20)
    E, g = pot.get_energy_gradient(x)
    fd = fd_grad(pot.get_energy, x)
    assert np.allclose(g * TWO_PI, fd, rtol=1e-6, atol=1e-8 * np.abs(fd).max())


def test_uwnu_mode_layout_first_dimension_fastest():
    m = O.uwnu_mode_index((4, 3))
    assert m[:5].tolist() == [[-2, -1], [-1, -1], [0, -1], [1, -1], [-2, 0]]


def test_target_grads_fd():
    # d phi / d S0 and d phi / d V by central differences of the loss itself
    rng = np.random.default_rng(12)
    m = 40
    rho = rng.normal(size=m) + 1j * rng.normal(size=m)
    Sk = rng.uniform(0.3, 2.0, m)
    Sk[::4] = 0.0
    V = rng.uniform(0, 1, m)
    norm = 3.0
    dS0, dV = O.target_grads(rho, Sk, V, norm)
    h = 1e-6
    for i in range(m):
        e = np.zeros(m)
        e[i] = h
        fv = (O._loss(rho, Sk, V + e, norm)[0] - O._loss(rho, Sk, V - e, norm)[0]) / (2 * h)
        assert abs(fv - dV[i]) <= 1e-6 * max(1.0, abs(dV[i]))
        if Sk[i] != 0:
            fs = (O._loss(rho, Sk + e, V, norm)[0] - O._loss(rho, Sk - e, V, norm)[0]) / (2 * h)
            assert abs(fs - dS0[i]) <= 1e-5 * max(1.0, abs(dS0[i]))
        else:
            # one-sided: the S0 == 0 branch (D = s - S0) is only taken at 0;
            # its derivative is that of (s - S0)^2 V
            s_i = abs(rho[i]) ** 2 / norm
            assert dS0[i] == pytest.approx(-2 * V[i] * s_i)
