"""Exact FP64 direct-sum oracle for the four structure-factor potentials.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Each function restates the arithmetic of the reference file:line it cites, with
every non-uniform FFT replaced by the exact sum it approximates (FINUFFT's
contract, Barnett, Magland & af Klinteberg 2019: type 1 f_k = sum_j c_j
e^{+-i k.x_j}; type 2 c_j = sum_k f_k e^{+-i k.x_j}; type 3 with arbitrary k).
So these functions are the reference path evaluated at eps -> 0.

Parity pinning (tests/test_oracle.py):
  kvec_size / mode_grid   closed form of nuwu.hpp:54-79 incl. truncation of
                          even sizes; square-lattice Bragg peaks land on the
                          mode indices the reference loop assigns.
  nuwu_weights            hand-computed radii cases (2D, 3D, 1D).
  type-1 / type-2 sums    separable sums == naive dense sums on tiny inputs;
                          N=1 |rho|=c; two points |rho|^2 = 2+2cos(k.dx);
                          square lattice rho = N on the reciprocal lattice, 0
                          elsewhere; Hermitian symmetry; adjointness
                          <f, T1 c> = <T2 f, c>.
  energies                lattice closed forms (E = count * N^2 etc.).
  gradients               central finite differences of the energies (NUwU,
                          NUwNU with c=1, UwNU x 1/(2pi)); the quirks of
                          NUwNU (no conj(c)) and UwU (2/sum I, no d(sum I))
                          are pinned by their exact-derivative relations.
"""
from __future__ import annotations

import math

import numpy as np

TWO_PI = 2.0 * math.pi


# ---------------------------------------------------------------------------
# conventions
# ---------------------------------------------------------------------------

def c_round(x):
    """C ``round``: half away from zero (numpy's np.round is half-to-even)."""
    x = np.asarray(x, dtype=np.float64)
    return np.copysign(np.floor(np.abs(x) + 0.5), x)


def kvec_size(m: int, d: int) -> int:
    """K' of ``calculate_Kvec`` (nuwu.hpp:54-79): the mode grid is (2K'+1)^d.

    1D uses size_t integer arithmetic, 2D/3D truncate sqrt/cbrt in double.
    """
    if d == 1:
        return (int(m) - 1) // 2
    if d == 2:
        return int((math.sqrt(m) - 1.0) / 2.0)
    if d == 3:
        return int((np.cbrt(float(m)) - 1.0) / 2.0)
    raise RuntimeError("Invalid dimension: %d" % d)


def mode_grid(kp: int, d: int) -> np.ndarray:
    """k vectors [M, d] in the reference's flat mode order (nuwu.hpp:277-285).

    Flat index i <-> kx = Kvec[i % n], ky = Kvec[(i % n^2) // n],
    kz = Kvec[i // n^2] with Kvec = -K'..K' (x fastest, FINUFFT CMCL order).
    """
    n = 2 * kp + 1
    kv = np.arange(-kp, kp + 1, dtype=np.int64)
    if d == 1:
        return kv.reshape(-1, 1)
    if d == 2:
        ky, kx = np.meshgrid(kv, kv, indexing="ij")
        return np.stack([kx.ravel(), ky.ravel()], axis=1)
    kz, ky, kx = np.meshgrid(kv, kv, kv, indexing="ij")
    return np.stack([kx.ravel(), ky.ravel(), kz.ravel()], axis=1)


def nuwu_weights(radii, d: int) -> np.ndarray:
    """Point weights c (nuwu.hpp:125-149; nuwnu.hpp:81-105).

    c_i = (pi r^2)^d in 2D, ((4/3) pi r^3)^d in 3D (the inner loop multiplies
    d times), 1 in 1D; then c *= N / sum(c).
    """
    r = np.asarray(radii, dtype=np.float64)
    c = np.ones_like(r)
    for _ in range(d):
        if d == 2:
            c = c * math.pi * r * r
        elif d == 3:
            c = c * math.pi * r * r * r * 4.0 / 3.0
    s = c.sum()
    return c * (len(c) / s)


def coords_shifted(p, L):
    """x = (p - round(p/L) + L) * 2pi/L, per component (nuwu.hpp:160,168,181;
    nuwnu.hpp:134,142,155,193,202,217).  Returns [N, d]."""
    L = np.asarray(L, dtype=np.float64)
    d = len(L)
    p = np.asarray(p, dtype=np.float64).reshape(-1, d)
    return (p - c_round(p / L) + L) * TWO_PI / L


def coords_centered(p, L):
    """x = (p - round(p/L)) * 2pi/L (nuwu.hpp:219,228,243).  Returns [N, d]."""
    L = np.asarray(L, dtype=np.float64)
    d = len(L)
    p = np.asarray(p, dtype=np.float64).reshape(-1, d)
    return (p - c_round(p / L)) * TWO_PI / L


# ---------------------------------------------------------------------------
# exact non-uniform sums
# ---------------------------------------------------------------------------

def _chunks(n, size):
    for s in range(0, n, size):
        yield s, min(n, s + size)


def nudft_dense(x, c, k, sign):
    """f_q = sum_j c_j exp(sign * i * k_q . x_j) by a dense loop over chunks.

    x [N, d], c [N] complex, k [M, d] real.  Brute force; any k."""
    x = np.asarray(x, dtype=np.float64)
    k = np.asarray(k, dtype=np.float64)
    c = np.asarray(c, dtype=np.complex128)
    out = np.zeros(k.shape[0], dtype=np.complex128)
    step = max(1, (1 << 22) // max(1, x.shape[0]))
    for a, b in _chunks(k.shape[0], step):
        ph = k[a:b] @ x.T  # [m, N]
        out[a:b] = np.exp(sign * 1j * ph) @ c
    return out


def nudft_dense_adj(x, f, k, sign):
    """u_j = sum_q f_q exp(sign * i * k_q . x_j) (type-2 / type-3 direction)."""
    x = np.asarray(x, dtype=np.float64)
    k = np.asarray(k, dtype=np.float64)
    f = np.asarray(f, dtype=np.complex128)
    out = np.zeros(x.shape[0], dtype=np.complex128)
    step = max(1, (1 << 22) // max(1, k.shape[0]))
    for a, b in _chunks(x.shape[0], step):
        ph = x[a:b] @ k.T  # [n, M]
        out[a:b] = np.exp(sign * 1j * ph) @ f
    return out


def _phases(x, kp):
    """e[j, t] = exp(i * (t - kp) * x_j) for t in 0..2kp, evaluated directly."""
    kv = np.arange(-kp, kp + 1, dtype=np.float64)
    return np.exp(1j * np.outer(x, kv))


def type1_grid(x, c, kp, sign):
    """rho in reference flat order for the (2kp+1)^d integer mode grid.

    Exact separable evaluation of sum_j c_j exp(sign i k.x_j); returns [M]."""
    x = np.asarray(x, dtype=np.float64)
    d = x.shape[1]
    c = np.asarray(c, dtype=np.complex128)
    n = 2 * kp + 1
    e = [_phases(x[:, a], kp) for a in range(d)]
    if sign < 0:
        e = [np.conj(t) for t in e]
    if d == 1:
        return c @ e[0]
    if d == 2:
        # rho[ky, kx] = sum_j c_j ey[j,ky] ex[j,kx]
        return ((c[:, None] * e[1]).T @ e[0]).ravel()
    out = np.zeros((n, n, n), dtype=np.complex128)
    step = max(1, (1 << 21) // (n * n))
    for a, b in _chunks(x.shape[0], step):
        A = (c[a:b, None, None] * e[2][a:b, :, None] * e[1][a:b, None, :])
        out += (A.reshape(b - a, n * n).T @ e[0][a:b]).reshape(n, n, n)
    return out.ravel()


def type2_grid(x, f, kp, sign):
    """u_j = sum_k f_k exp(sign i k.x_j) over the integer mode grid (f in
    reference flat order); exact separable evaluation; returns [N]."""
    x = np.asarray(x, dtype=np.float64)
    d = x.shape[1]
    n = 2 * kp + 1
    f = np.asarray(f, dtype=np.complex128)
    e = [_phases(x[:, a], kp) for a in range(d)]
    if sign < 0:
        e = [np.conj(t) for t in e]
    if d == 1:
        return e[0] @ f
    if d == 2:
        F = f.reshape(n, n)  # [ky, kx]
        T = e[0] @ F.T  # [N, ky]
        return np.sum(T * e[1], axis=1)
    F = f.reshape(n, n, n)  # [kz, ky, kx]
    out = np.zeros(x.shape[0], dtype=np.complex128)
    step = max(1, (1 << 21) // (n * n))
    for a, b in _chunks(x.shape[0], step):
        T = e[0][a:b] @ F.reshape(n * n, n).T  # [nb, kz*ky]
        T = T.reshape(b - a, n, n)
        out[a:b] = np.einsum("jzy,jz,jy->j", T, e[2][a:b], e[1][a:b])
    return out


# ---------------------------------------------------------------------------
# shared loss (the per-mode loop of every potential)
# ---------------------------------------------------------------------------

def _loss(rho, Sk, V, norm):
    """s = |rho|^2/norm; D = (s - S0)[/S0 where S0 != 0]; phi = sum V D^2;
    g = V*D[/S0 where S0 != 0]  (nuwu.hpp:260-276, nuwnu.hpp:235-251,
    uwu.hpp:166-182, uwnu.hpp:131-147).  Returns (phi, g) with the
    per-mode gradient factor before the potential's own constant."""
    Sk = np.asarray(Sk, dtype=np.float64)
    V = np.asarray(V, dtype=np.float64)
    a = np.abs(rho)
    D = a * a / norm - Sk
    nz = Sk != 0
    D = np.where(nz, D / np.where(nz, Sk, 1.0), D)
    phi = float(np.sum(V * D * D))
    g = V * D
    g = np.where(nz, g / np.where(nz, Sk, 1.0), g)
    return phi, g


def target_grads(rho, Sk, V, norm):
    """d phi / d S0 and d phi / d V of phi = sum V D^2 (the loss of
    nuwu.hpp:260-276 as a function of its targets): dphi/dV = D^2;
    dphi/dS0 = -2 V D s / S0^2 where S0 != 0 (D = s/S0 - 1), else -2 V D
    (D = s - S0), s = |rho|^2 / norm."""
    Sk = np.asarray(Sk, dtype=np.float64)
    V = np.asarray(V, dtype=np.float64)
    s = np.abs(rho) ** 2 / norm
    nz = Sk != 0
    Ssafe = np.where(nz, Sk, 1.0)
    D = np.where(nz, (s - Sk) / Ssafe, s - Sk)
    dS0 = np.where(nz, -2.0 * V * D * s / Ssafe ** 2, -2.0 * V * D)
    return dS0, D * D


# ---------------------------------------------------------------------------
# NUwU  (source/fresco/nuwu.hpp)
# ---------------------------------------------------------------------------

class NUwU:
    """Direct-sum restatement of fresco::NUwU (nuwu.hpp:17-318)."""

    def __init__(self, radii, K, Sk, V, L, eps=1e-6):
        self.L = np.asarray(L, dtype=np.float64)
        self.d = len(self.L)
        self.Sk = np.asarray(Sk, dtype=np.float64).ravel()
        self.V = np.asarray(V, dtype=np.float64).ravel()
        assert len(self.Sk) == len(self.V)
        self.N = len(radii)
        self.kp = kvec_size(len(self.V), self.d)
        self.c = nuwu_weights(radii, self.d)
        self.k = mode_grid(self.kp, self.d)

    def rho(self, p, shifted=False):
        x = coords_shifted(p, self.L) if shifted else coords_centered(p, self.L)
        return type1_grid(x, self.c, self.kp, +1)

    def get_energy(self, p):
        """nuwu.hpp:152-208 (coordinate map with +L, type-1 +1, loss)."""
        rho = self.rho(p, shifted=True)
        phi, _ = _loss(rho, self.Sk[: len(rho)], self.V[: len(rho)], self.N)
        return phi

    def get_energy_gradient(self, p):
        """nuwu.hpp:210-317.  grad[d*j+a] = Re(cx_a[j] conj(c_j)) with
        cx_a = type-2 (-1) of f_a = -i k_a (4/N) V D rho [/S0]."""
        x = coords_centered(p, self.L)
        rho = type1_grid(x, self.c, self.kp, +1)
        M = len(rho)
        phi, g = _loss(rho, self.Sk[:M], self.V[:M], self.N)
        factor = (4.0 / self.N) * g * rho
        grad = np.zeros((self.N, self.d))
        for a in range(self.d):
            fa = self.k[:, a] * factor * (-1j)
            cx = type2_grid(x, fa, self.kp, -1)
            grad[:, a] = np.real(cx * np.conj(self.c))
        return phi, grad.ravel()


# ---------------------------------------------------------------------------
# NUwNU  (source/fresco/nuwnu.hpp)
# ---------------------------------------------------------------------------

class NUwNU:
    """Direct-sum restatement of fresco::NUwNU (nuwnu.hpp:17-291).

    k vectors are used unscaled (nuwnu.hpp:47-76) while x is scaled by 2pi/L
    with the +L shift in both paths (nuwnu.hpp:134,193)."""

    def __init__(self, radii, K, Sk, V, L, eps=1e-6):
        self.L = np.asarray(L, dtype=np.float64)
        self.d = len(self.L)
        self.N = len(radii)
        Kf = np.asarray(K, dtype=np.float64).ravel()
        self.nk = Kf.size // self.d
        self.k = Kf[: self.nk * self.d].reshape(self.nk, self.d)
        self.Sk = np.asarray(Sk, dtype=np.float64).ravel()
        self.V = np.asarray(V, dtype=np.float64).ravel()
        self.c = nuwu_weights(radii, self.d)

    def rho(self, p):
        x = coords_shifted(p, self.L)
        return nudft_dense(x, self.c, self.k, +1)

    def get_energy(self, p):
        rho = self.rho(p)
        phi, _ = _loss(rho, self.Sk[: self.nk], self.V[: self.nk], self.N)
        return phi

    def get_energy_gradient(self, p):
        """nuwnu.hpp:184-290: f_a = k_a (-4i/N) V D rho [/S0];
        grad = Re(type-3 (-1) of f_a) with no conj(c) (nuwnu.hpp:269)."""
        x = coords_shifted(p, self.L)
        rho = nudft_dense(x, self.c, self.k, +1)
        phi, g = _loss(rho, self.Sk[: self.nk], self.V[: self.nk], self.N)
        factor = (-4.0j / self.N) * g * rho
        grad = np.zeros((self.N, self.d))
        for a in range(self.d):
            cx = nudft_dense_adj(x, self.k[:, a] * factor, self.k, -1)
            grad[:, a] = np.real(cx)
        return phi, grad.ravel()


# ---------------------------------------------------------------------------
# UwU  (source/fresco/uwu.hpp, fresco/potentials/uwu.pyx)
# ---------------------------------------------------------------------------

def dft_grid(a, sign):
    """Unnormalised multi-dimensional DFT with exponent sign (FFTW convention:
    sign +1 = sum a_m e^{+2 pi i k.m / N}).  Row-major, last axis fastest."""
    a = np.asarray(a, dtype=np.complex128)
    if sign > 0:
        return np.fft.ifftn(a) * a.size
    return np.fft.fftn(a)


class UwU:
    """Restatement of fresco::UwU (uwu.hpp:17-195) with the wrapper's
    ifftshift of Sk and V (uwu.pyx:14-15).  The racy ``rhotot +=``
    (uwu.hpp:128-133) is evaluated as the exact sum."""

    def __init__(self, N, K, Sk, V, L, eps=1e-6):
        self.shape = tuple(int(v) for v in np.asarray(N).ravel())
        self.Sk = np.fft.ifftshift(np.asarray(Sk, dtype=np.float64)).ravel()
        self.V = np.fft.ifftshift(np.asarray(V, dtype=np.float64)).ravel()
        assert len(self.Sk) == len(self.V)
        self.L = np.asarray(L, dtype=np.float64)
        self.d = len(self.L)

    def get_energy(self, x):
        x = np.asarray(x, dtype=np.float64).ravel()
        rho = dft_grid(x.reshape(self.shape), +1).ravel()
        phi, _ = _loss(rho, self.Sk, self.V, x.sum())
        return phi

    def get_energy_gradient(self, x):
        """uwu.hpp:152-194: f = (2/sum I) V D rho [/S0];
        grad = Re(FFT_{-1}(f))."""
        x = np.asarray(x, dtype=np.float64).ravel()
        rhotot = x.sum()
        rho = dft_grid(x.reshape(self.shape), +1).ravel()
        phi, g = _loss(rho, self.Sk, self.V, rhotot)
        f = (2.0 / rhotot) * g * rho
        grad = np.real(dft_grid(f.reshape(self.shape), -1)).ravel()
        return phi, grad


# ---------------------------------------------------------------------------
# UwNU  (source/fresco/uwnu.hpp)
# ---------------------------------------------------------------------------

def uwnu_mode_index(shape):
    """Integer mode coordinates [P, d] of the grid values as FINUFFT modes:
    flat index j = i0 + N0*(i1 + N1*i2) (first dimension fastest, n_modes =
    (N[0], N[1], N[2]) at uwnu.hpp:44-82), m_a = i_a - floor(N_a/2)."""
    d = len(shape)
    axes = [np.arange(n, dtype=np.int64) - (n // 2) for n in shape]
    grids = np.meshgrid(*axes[::-1], indexing="ij")  # slowest first
    return np.stack([g.ravel() for g in grids[::-1]], axis=1)


class UwNU:
    """Restatement of fresco::UwNU (uwnu.hpp:17-159): grid values are the
    NUFFT modes, k scaled by 2pi/L (uwnu.hpp:48,58,73), rhotot = 1,
    type-2 +1 for rho, f = 4 V D rho [/S0], grad = Re(type-1 (-1) f)/(2pi)."""

    def __init__(self, N, K, Sk, V, L, eps=1e-6):
        self.shape = tuple(int(v) for v in np.asarray(N).ravel())
        self.L = np.asarray(L, dtype=np.float64)
        self.d = len(self.L)
        self.Sk = np.asarray(Sk, dtype=np.float64).ravel()
        self.V = np.asarray(V, dtype=np.float64).ravel()
        self.nk = len(self.Sk)
        Kf = np.asarray(K, dtype=np.float64).ravel()
        self.k = Kf[: self.nk * self.d].reshape(self.nk, self.d) * TWO_PI / self.L
        self.m = uwnu_mode_index(self.shape)

    def rho(self, x):
        x = np.asarray(x, dtype=np.float64).ravel()
        return nudft_dense(self.m.astype(np.float64), x.astype(np.complex128), self.k, +1)

    def get_energy(self, x):
        rho = self.rho(x)
        phi, _ = _loss(rho, self.Sk, self.V, 1.0)
        return phi

    def get_energy_gradient(self, x):
        rho = self.rho(x)
        phi, g = _loss(rho, self.Sk, self.V, 1.0)
        f = 4.0 * g * rho
        c = nudft_dense_adj(self.m.astype(np.float64), f, self.k, -1)
        return phi, np.real(c) / TWO_PI
