"""Pins the CPU NUFFT restatement (oracle/nufft_cpu.py) against the exact
direct sums of oracle/sf_direct.py (CPU only)."""
import numpy as np
import pytest

from oracle import nufft_cpu as C
from oracle import sf_direct as D


@pytest.fixture(scope="module", autouse=True)
def built():
    import subprocess
    import os

    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    subprocess.run(["make", "-s", "-C", here], check=True)


@pytest.mark.parametrize("d,kp,N", [(1, 20, 300), (2, 10, 500), (3, 5, 400)])
@pytest.mark.parametrize("eps", [1e-6, 1e-12])
def test_type1_type2_vs_direct(d, kp, N, eps):
    rng = np.random.default_rng(d)
    x = rng.uniform(-7, 7, (N, d))
    c = rng.normal(size=N) + 1j * rng.normal(size=N)
    g = C.NufftGrid(d, kp, eps)
    r = g.type1(x, c)
    r0 = D.type1_grid(x, c, kp, +1)
    assert np.linalg.norm(r - r0) / np.linalg.norm(r0) < 10 * eps
    f = rng.normal(size=len(r0)) + 1j * rng.normal(size=len(r0))
    u = g.type2(x, f)
    u0 = D.type2_grid(x, f, kp, -1)
    assert np.linalg.norm(u - u0) / np.linalg.norm(u0) < 10 * eps


@pytest.mark.parametrize("d", [2, 3])
def test_nuwu_energy_gradient_vs_direct(d):
    rng = np.random.default_rng(5 + d)
    N, n = 300, (15 if d == 2 else 9)
    M = n ** d
    radii = rng.uniform(0.5, 1.5, N)
    Sk = rng.uniform(0, 2, M)
    Sk[::3] = 0
    V = rng.uniform(0, 1, M)
    p = rng.uniform(-0.2, 1.2, N * d)
    a = C.NUwU(radii, 0, Sk, V, np.ones(d), eps=1e-12)
    b = D.NUwU(radii, 0, Sk, V, np.ones(d))
    E, g = a.get_energy_gradient(p)
    E0, g0 = b.get_energy_gradient(p)
    assert abs(E - E0) / E0 < 1e-10
    assert np.linalg.norm(g - g0) / np.linalg.norm(g0) < 1e-10
    assert abs(a.get_energy(p) - b.get_energy(p)) / E0 < 1e-10
