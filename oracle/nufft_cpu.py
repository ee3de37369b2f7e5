"""CPU FP64 restatement of the reference NUwU path with a real NUFFT.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the timed CPU baseline
("CPU restatement of the reference path") and a checker at sizes where the
exact direct sum of ``sf_direct`` is too slow.

Structure follows source/fresco/nuwu.hpp:210-317 exactly as the reference
runs it: a complex type-1 NUFFT (+1) of the weights (spread on a complex fine
grid, FFT, deconvolve / crop), the per-mode loss, then d separate type-2
NUFFTs (-1) of f_a = -i k_a factor (pad / deconvolve, FFT, interpolate) and
grad = Re(cx conj(c)).  The NUFFT is the exponential-of-semicircle method of
Barnett, Magland & af Klinteberg (2019) with the same width/beta rule and
upsampling as FINUFFT's defaults; spreading / interpolation run in
oracle/nufft_cpu.c (OpenMP), the FFTs in scipy.fft (all host threads).

Pinning: tests/test_oracle_nufft.py compares it with the exact direct sums of
``sf_direct`` (type 1, type 2 and the full energy / gradient).
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import scipy.fft

from . import sf_direct as D

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libnufft_cpu.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/libnufft_cpu.so missing: run `make -C oracle`")
        lib = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        for name in ("nufft_spread", "nufft_interp"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int, ctypes.c_int64, dp, dp, dp, dp, ip, ctypes.c_int, ctypes.c_double, dp]
        lib.nufft_threads.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def threads():
    return int(_lib().nufft_threads())


def es_params(eps, sigma=2.0):
    """Kernel width and beta (Barnett et al. 2019, the rule FINUFFT uses)."""
    if sigma == 2.0:
        w = max(2, int(math.ceil(-math.log10(eps / 10.0))))
        bw = {2: 2.20, 3: 2.26, 4: 2.38}.get(w, 2.30)
    else:
        w = max(2, int(math.ceil(-math.log(eps) / (math.pi * math.sqrt(1.0 - 1.0 / sigma)))))
        bw = 0.97 * math.pi * (1.0 - 1.0 / (2.0 * sigma))
    return w, bw * w


def next_smooth_even(target, w):
    n = max(int(target), 2 * w)
    n += n % 2
    while True:
        m = n
        for f in (2, 3, 5):
            while m % f == 0:
                m //= f
        if m == 1:
            return n
        n += 2


def deconv(w, beta, nf, kmax):
    """Psi(k) = int phi(2t/w) cos(2 pi k t / nf) dt, k = 0..kmax."""
    z, wt = np.polynomial.legendre.leggauss(4 * w + 64)
    ph = np.exp(beta * (np.sqrt(1 - z * z) - 1))
    k = np.arange(kmax + 1)
    return 0.5 * w * (np.cos(np.pi * np.outer(k, z) * w / nf) @ (wt * ph))


class NufftGrid:
    """Type-1 (+1) / type-2 (-1) NUFFT between points and the (2kp+1)^d modes."""

    def __init__(self, d, kp, eps, sigma=2.0, workers=None):
        self.d = d
        self.kp = kp
        self.n = 2 * kp + 1
        self.w, self.beta = es_params(eps, sigma)
        self.nf = [next_smooth_even(math.ceil(sigma * self.n), self.w)] * d
        self.psi = [deconv(self.w, self.beta, nf, kp) for nf in self.nf]
        self.workers = workers or threads()
        kv = np.arange(-kp, kp + 1)
        self.idx = [np.mod(kv, nf) for nf in self.nf]
        ip = [1.0 / p[np.abs(kv)] for p in self.psi]
        # deconvolution weights on the mode box, array order [kz, ky, kx]
        w = ip[0]
        for a in range(1, d):
            w = np.multiply.outer(ip[a], w)
        self.ipsi = w

    def _xs(self, x):
        cols = [np.ascontiguousarray(x[:, a]) for a in range(self.d)]
        while len(cols) < 3:
            cols.append(cols[0])
        return cols

    def _box(self):
        return np.ix_(*[self.idx[a] for a in range(self.d - 1, -1, -1)])

    def type1(self, x, c):
        """rho_k = sum_j c_j e^{+i k.x_j}, reference flat order."""
        lib = _lib()
        n = x.shape[0]
        shape = self.nf[::-1]
        grid = np.zeros(shape + [2])
        cc = np.ascontiguousarray(np.stack([np.real(c), np.imag(c)], -1).astype(np.float64))
        xs = self._xs(x)
        nf = (ctypes.c_int * 3)(*(self.nf + [1] * (3 - self.d)))
        dp = ctypes.POINTER(ctypes.c_double)
        rc = lib.nufft_spread(self.d, n, *[a.ctypes.data_as(dp) for a in xs], cc.ctypes.data_as(dp), nf, self.w,
                              self.beta, grid.ctypes.data_as(dp))
        assert rc == 0
        g = grid[..., 0] + 1j * grid[..., 1]
        B = scipy.fft.ifftn(g, workers=self.workers) * g.size  # sum b_l e^{+i k l h}
        return (B[self._box()] * self.ipsi).ravel()

    def type2(self, x, f):
        """u_j = sum_k f_k e^{-i k.x_j} for f in reference flat order."""
        lib = _lib()
        shape = self.nf[::-1]
        F = np.zeros(shape, dtype=np.complex128)
        F[self._box()] = f.reshape([self.n] * self.d) * self.ipsi
        b = scipy.fft.fftn(F, workers=self.workers)  # sum_k F_k e^{-i k l h}
        bb = np.ascontiguousarray(np.stack([b.real, b.imag], -1))
        out = np.zeros((x.shape[0], 2))
        xs = self._xs(x)
        nf = (ctypes.c_int * 3)(*(self.nf + [1] * (3 - self.d)))
        dp = ctypes.POINTER(ctypes.c_double)
        rc = lib.nufft_interp(self.d, x.shape[0], *[a.ctypes.data_as(dp) for a in xs], bb.ctypes.data_as(dp), nf,
                              self.w, self.beta, out.ctypes.data_as(dp))
        assert rc == 0
        return out[:, 0] + 1j * out[:, 1]


class NUwU:
    """CPU FP64 NUFFT restatement of fresco::NUwU (nuwu.hpp:17-318)."""

    def __init__(self, radii, K, Sk, V, L, eps=1e-6, sigma=2.0, workers=None):
        self.L = np.asarray(L, dtype=np.float64)
        self.d = len(self.L)
        self.Sk = np.asarray(Sk, dtype=np.float64).ravel()
        self.V = np.asarray(V, dtype=np.float64).ravel()
        self.N = len(radii)
        self.kp = D.kvec_size(len(self.V), self.d)
        self.c = D.nuwu_weights(radii, self.d)
        self.k = D.mode_grid(self.kp, self.d)
        self.nufft = NufftGrid(self.d, self.kp, eps, sigma, workers)

    def get_energy(self, p):
        x = D.coords_shifted(p, self.L)
        rho = self.nufft.type1(x, self.c.astype(np.complex128))
        phi, _ = D._loss(rho, self.Sk, self.V, self.N)
        return phi

    def get_energy_gradient(self, p):
        x = D.coords_centered(p, self.L)
        rho = self.nufft.type1(x, self.c.astype(np.complex128))
        phi, g = D._loss(rho, self.Sk, self.V, self.N)
        factor = (4.0 / self.N) * g * rho
        grad = np.zeros((self.N, self.d))
        for a in range(self.d):
            cx = self.nufft.type2(x, self.k[:, a] * factor * (-1j))
            grad[:, a] = np.real(cx * np.conj(self.c))
        return phi, grad.ravel()
