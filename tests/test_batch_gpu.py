"""Batched NUwU (many independent patterns in one evaluation) against the
oracle's exact sums per pattern, and the ensemble configuration: 256 2D
patterns, N log-uniform in [2^12, 2^18], FP32 against FP64 per pattern.

Tolerances: FP64 (eps 1e-12) 1e-9 per pattern; FP32 (eps 1e-6) 1e-5.
"""
import time

import numpy as np
import pytest

from oracle import sf_direct as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def P():
    import fresco_amd.potentials as P

    return P


def small_patterns(d, seed):
    rng = np.random.default_rng(seed)
    pats, xs = [], []
    for N, n in ((60, 9), (300, 17), (150, 13), (500, 21), (90, 11)):
        M = n ** d
        Sk = rng.uniform(0.3, 2, M)
        Sk[::3] = 0
        V = rng.uniform(0, 1, M)
        V[::5] = 0
        pats.append((rng.uniform(0.6, 1.4, N), Sk, V))
        xs.append(rng.uniform(-0.2, 1.2, N * d))
    return pats, np.concatenate(xs)


@pytest.mark.parametrize("d", [1, 2])
@pytest.mark.parametrize("prec,eps,tol", [("fp64", 1e-12, 1e-9), ("fp32", 1e-6, 1e-5)])
def test_batch_vs_exact_sums_per_pattern(P, d, prec, eps, tol):
    pats, x = small_patterns(d, seed=d)
    L = np.ones(d) * (1.0 if d == 2 else 1.0)
    pot = P.NUwUBatch(pats, L, eps, precision=prec)
    E, g = pot.get_energy_gradient(x)
    Ep = pot.pattern_energies()
    gs = pot.split(g)
    xs = pot.split(x)
    tot = 0.0
    for p, (r, Sk, V) in enumerate(pats):
        E0, g0 = O.NUwU(r, 0, Sk, V, L).get_energy_gradient(xs[p])
        tot += E0
        assert abs(Ep[p] - E0) / abs(E0) < tol
        assert rel(gs[p], g0) < tol
    assert abs(E - tot) / tot < tol
    assert abs(pot.get_energy(x) - tot) / tot < tol
    E2, g2 = pot.get_energy_gradient(x)
    assert E2 == E and np.array_equal(g2, g)


def ensemble(count, seed0=1000):
    """The ensemble configuration: N = round(2^(12 + 6 u)), n = 2 round(32
    (N / 2^12)^(1/3)) + 1 modes per dimension, V = 1 on 0 < |k| <= sqrt(N)/8,
    S0 = 0, coordinates uniform in the unit box (seed 1000 + i)."""
    rng = np.random.default_rng(7)
    u = rng.uniform(0, 1, count)
    pats, xs = [], []
    for i in range(count):
        N = int(round(2 ** (12 + 6 * u[i])))
        n = 2 * int(round(32 * (N / 4096) ** (1 / 3))) + 1
        kp = (n - 1) // 2
        k = O.mode_grid(kp, 2)
        kk = np.linalg.norm(k, axis=1)
        V = ((kk > 0) & (kk <= np.sqrt(N) / 8)).astype(float)
        pats.append((np.ones(N), np.zeros_like(V), V))
        xs.append(np.random.default_rng(seed0 + i).uniform(0, 1, 2 * N))
    return pats, xs


def test_ensemble_256_fp32_vs_fp64(P):
    import torch

    pats, xs = ensemble(256)
    L = np.ones(2)
    x = torch.tensor(np.concatenate(xs), device="cuda")
    b32 = P.NUwUBatch(pats, L, 1e-6, precision="fp32")
    b64 = P.NUwUBatch(pats, L, 1e-9, precision="fp64")
    g32 = torch.empty_like(x)
    g64 = torch.empty_like(x)
    for _ in range(2):
        b32.energy_gradient_device(x, g32)
    torch.cuda.synchronize()
    t = time.perf_counter()
    reps = 10
    for _ in range(reps):
        b32.energy_gradient_device(x, g32)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    b64.energy_gradient_device(x, g64)
    E32, E64 = b32.pattern_energies(), b64.pattern_energies()
    re = np.abs(E32 - E64) / np.abs(E64)
    a, b = g32.cpu().numpy(), g64.cpu().numpy()
    off = b32.offsets
    rg = np.array([rel(a[off[p]:off[p + 1]], b[off[p]:off[p + 1]]) for p in range(256)])
    print("ensemble: %d patterns, %d points, %.3f ms per batched E+grad, %.0f patterns/s; worst rel err "
          "energy %.2e gradient %.2e" % (256, int(b32.npts.sum()), dt * 1e3, 256 / dt, re.max(), rg.max()),
          flush=True)
    assert re.max() < 1e-5
    assert rg.max() < 1e-5
    # the batch agrees with single-pattern plans (smallest and largest pattern)
    from fresco_amd.potentials import NUwU

    for p in (int(np.argmin(b32.npts)), int(np.argmax(b32.npts))):
        r, Sk, V = pats[p]
        E1, g1 = NUwU(r, 0, Sk, V, L, 1e-9, precision="fp64-adjoint").get_energy_gradient(xs[p])
        assert abs(E64[p] - E1) / abs(E1) < 1e-8
        assert rel(b[off[p]:off[p + 1]], g1) < 1e-8


@pytest.mark.parametrize("case", ["small1d", "small2d", "ensemble24"])
def test_segmented_direct_binning_equals_radix_sort(P, case, monkeypatch):
    """Batched plans bin each pattern's points by one counting pass over tiles
    inside the pattern (counts, per-pattern prefixes over tiles and bins,
    ballot-ranked placement); the segmented radix sort
    (FRESCO_DIRECT_BINNING=0) gives the same stable order, so FP32 energies,
    pattern energies and gradients agree bitwise.  The ensemble case has
    patterns of 2^12..2^18 points (up to 256 tiles per pattern)."""
    if case == "ensemble24":
        pats, xs = ensemble(24, seed0=2000)
        x, L = np.concatenate(xs), np.ones(2)
    else:
        d = 1 if case == "small1d" else 2
        pats, x = small_patterns(d, seed=10 + d)
        L = np.ones(d)
    res = []
    for env in ("1", "0"):
        monkeypatch.setenv("FRESCO_DIRECT_BINNING", env)
        pot = P.NUwUBatch(pats, L, 1e-6, precision="fp32")
        E, g = pot.get_energy_gradient(x)
        res.append((E, g, pot.pattern_energies()))
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])
    assert np.array_equal(res[0][2], res[1][2])


def test_batch_lbfgs_lowers_every_pattern(P):
    from fresco_amd.optimize import LBFGS_CPP

    pats, xs = ensemble(8, seed0=2000)
    pot = P.NUwUBatch(pats, np.ones(2), 1e-6, precision="fp32")
    x0 = np.concatenate(xs)
    pot.get_energy(x0)
    E0 = pot.pattern_energies()
    res = LBFGS_CPP(x0, pot, M=10, nsteps=200, tol=0.0).run(200)
    pot.get_energy(res.coords)
    E1 = pot.pattern_energies()
    print("batch LBFGS pattern energies", E0, "->", E1, flush=True)
    assert np.all(E1 < E0)
