"""GPU parity of the device optimizers (LBFGS_CPP, ModifiedFireCPP) against the
optimizer oracle and the reference's own lbfgs.cpp / modified_fire.cpp
compiled from its sources (oracle/_ref), plus the Python-facing behaviour of
fresco/optimize (events, isave checkpoints, Result, reset)."""
import os

import numpy as np
import pytest

from oracle import optim as OP
from oracle import ref_opt as RO
from oracle import sf_direct as D

pytestmark = pytest.mark.gpu


class Harmonic:
    """E = 1/2 |x|^2 (cpp_tests/source/test_base_potential.cpp:11-25)."""

    def get_energy_gradient(self, x):
        x = np.asarray(x)
        return 0.5 * float(np.dot(x, x)), x.copy()

    def get_energy(self, x):
        return 0.5 * float(np.dot(x, x))


def nuwu_case(seed=0, N=48, n=13, d=2):
    rng = np.random.default_rng(seed)
    M = n ** d
    k = D.mode_grid(n // 2, d)
    kk = np.linalg.norm(k, axis=1)
    V = ((kk > 0) & (kk <= n // 3)).astype(float)
    Sk = np.zeros(M)
    return np.ones(N), Sk, V, rng.uniform(0, 1, N * d)


@pytest.fixture(scope="module")
def mods():
    from fresco_amd.optimize import LBFGS_CPP, ModifiedFireCPP
    from fresco_amd.potentials import CombinedPotential, NUwU

    return LBFGS_CPP, ModifiedFireCPP, NUwU, CombinedPotential


@pytest.mark.parametrize("M", [4, 10])
def test_lbfgs_nuwu_tracks_oracle(mods, M):
    LBFGS_CPP, _, NUwU, _ = mods
    radii, Sk, V, x0 = nuwu_case(M)
    pot = NUwU(radii, 0, Sk, V, np.ones(2), 1e-13, precision="fp64")
    ref_pot = D.NUwU(radii, 0, Sk, V, np.ones(2))
    opt = LBFGS_CPP(x0, pot, tol=1e-14, M=M)
    ora = OP.LBFGS(x0, ref_pot, tol=1e-14, M=M)
    ora.initialize_func_gradient()
    for it in range(25):
        r = opt.one_iteration()
        ora.one_iteration()
        print("iter", it, r.energy, ora.f, flush=True)
        assert r.nfev == ora.nfev and r.nsteps == ora.iter
        assert np.allclose(r.coords, ora.x, rtol=0, atol=1e-8)
        assert abs(r.energy - ora.f) <= 1e-8 * abs(ora.f) + 1e-12
        assert abs(r["H0"] - ora.H0) <= 1e-5 * abs(ora.H0)


def test_lbfgs_python_potential_vs_reference_binary(mods):
    LBFGS_CPP = mods[0]
    if not RO.available():
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(3)
    x0 = rng.normal(size=40)
    A = rng.normal(size=(40, 40))
    A = A @ A.T / 40 + np.eye(40)

    class Quad:
        def get_energy_gradient(self, x):
            return 0.5 * float(x @ A @ x), A @ x

    opt = LBFGS_CPP(x0, Quad(), tol=1e-10, M=5)
    ref = RO.LBFGS(x0, Quad(), tol=1e-10, M=5, max_f_rise=1e-4)
    for it in range(30):
        r = opt.one_iteration()
        ref.one_iteration()
        s = ref.state()
        assert r.nfev == s["nfev"]
        assert np.allclose(r.coords, s["x"], rtol=1e-9, atol=1e-12), it
        assert np.isclose(r["H0"], s["H0"], rtol=1e-8)


def test_fire_tracks_reference_binary(mods):
    _, ModifiedFireCPP, NUwU, _ = mods
    if not RO.available():
        pytest.skip("oracle/_ref not built")
    radii, Sk, V, x0 = nuwu_case(5, N=40, n=11)
    pot = NUwU(radii, 0, Sk, V, np.ones(2), 1e-13, precision="fp64")
    ref_pot = D.NUwU(radii, 0, Sk, V, np.ones(2))
    opt = ModifiedFireCPP(x0, pot, dtstart=0.001, dtmax=0.01, maxstep=0.01, tol=1e-12)
    ref = RO.ModifiedFire(x0, ref_pot, dtstart=0.001, dtmax=0.01, maxstep=0.01, tol=1e-12)
    opt.stop_criterion_satisfied()
    ref.stop_criterion_satisfied()
    for it in range(40):
        r = opt.one_iteration()
        ref.one_iteration()
        s = ref.state()
        assert r.nfev == s["nfev"] and r.nsteps == s["niter"]
        assert np.allclose(r.coords, s["x"], rtol=0, atol=1e-9), it
        assert abs(r.energy - s["f"]) <= 1e-8 * abs(s["f"])


def _c1_case():
    """2D N = 4096, 65^2 modes, stealthy 0 < |k| <= 16, S0 = 0, V = 1."""
    N, kp = 4096, 32
    k = D.mode_grid(kp, 2)
    kk = np.linalg.norm(k, axis=1)
    V = ((kk > 0) & (kk <= 16)).astype(float)
    return np.ones(N), np.zeros_like(V), V, np.random.default_rng(0).uniform(0, 1, 2 * N)


@pytest.mark.parametrize("case", ["n48", "n4096"])
def test_reference_lbfgs_drives_gpu_potential(mods, case):
    """The reference's own LBFGS (source/lbfgs.cpp, compiled from its source)
    runs the GPU NUwU through include/fresco_hip_potential.hpp
    (fresco::HipPotential, the BasePotential the reference's optimizers call
    at optimizer.hpp:208-215), and tracks the device LBFGS on the same
    potential for 25 iterations: same line-search decisions, positions to
    1e-8."""
    LBFGS_CPP, _, NUwU, _ = mods
    if not RO.hip_available():
        pytest.skip("oracle/_ref/libref_hip_opt.so not built (needs the reference sources)")
    if case == "n48":
        radii, Sk, V, x0 = nuwu_case(3)
        M = 4
    else:
        radii, Sk, V, x0 = _c1_case()
        M = 10
    pot = NUwU(radii, 0, Sk, V, np.ones(2), 1e-12, precision="fp64")
    ref = RO.HipLBFGS(pot._h, x0, tol=0.0, M=M)
    dev = LBFGS_CPP(x0, pot, tol=0.0, M=M)
    for it in range(1, 26):
        ref.one_iteration()
        r = dev.one_iteration()
        s = ref.state()
        if it % 5 == 0:
            print(case, "it", it, "E reference", s["f"], "device", r.energy, "max|dx| %.2e"
                  % np.max(np.abs(r.coords - s["x"])), flush=True)
        assert s["nfev"] == r.nfev and s["niter"] == r.nsteps == it
        assert np.allclose(r.coords, s["x"], rtol=0, atol=1e-8), it
        assert abs(r.energy - s["f"]) <= 1e-8 * abs(s["f"]) + 1e-14, it
        assert abs(r.rms - s["rms"]) <= 1e-8 * s["rms"], it


def test_reference_fire_drives_gpu_potential(mods):
    """The reference's MODIFIED_FIRE (modified_fire.hpp / .cpp) over the GPU
    NUwU through fresco::HipPotential, against the device FIRE on the same
    potential, iteration by iteration."""
    _, ModifiedFireCPP, NUwU, _ = mods
    if not RO.hip_available():
        pytest.skip("oracle/_ref/libref_hip_opt.so not built (needs the reference sources)")
    radii, Sk, V, x0 = nuwu_case(5, N=40, n=11)
    pot = NUwU(radii, 0, Sk, V, np.ones(2), 1e-13, precision="fp64")
    ref = RO.HipModifiedFire(pot._h, x0, dtstart=0.001, dtmax=0.01, maxstep=0.01, tol=1e-12)
    dev = ModifiedFireCPP(x0, pot, dtstart=0.001, dtmax=0.01, maxstep=0.01, tol=1e-12)
    dev.stop_criterion_satisfied()
    ref.stop_criterion_satisfied()  # both evaluate x0 first (optimizer.hpp:195-201)
    for it in range(1, 41):
        ref.one_iteration()
        r = dev.one_iteration()
        s = ref.state()
        assert s["nfev"] == r.nfev and s["niter"] == r.nsteps
        assert np.allclose(r.coords, s["x"], rtol=0, atol=1e-9), it
        assert abs(r.energy - s["f"]) <= 1e-9 * abs(s["f"]), it


def test_lbfgs_harmonic_converges_and_result_protocol(mods):
    LBFGS_CPP = mods[0]
    opt = LBFGS_CPP(np.array([1.0, 1.5]), Harmonic(), tol=1e-10)
    res = opt.run(200)
    assert res.success and np.allclose(res.coords, 0, atol=1e-9)
    coords, energy, rms, nfev, same = res
    assert energy == res.energy and rms == res[2] and nfev == res[3] and same is res or same == res
    assert res[:2][1] == res.energy


def test_events_checkpoints_reset(mods, tmp_path):
    LBFGS_CPP, _, NUwU, _ = mods
    radii, Sk, V, x0 = nuwu_case(7)
    pot = NUwU(radii, 0, Sk, V, np.ones(2), 1e-10, precision="fp64")
    seen = []
    opt = LBFGS_CPP(x0, pot, events=[lambda coords, energy, rms: seen.append(energy)])
    res = opt.run(5)
    assert len(seen) == 5 and res.nsteps == 5
    opt2 = LBFGS_CPP(x0, pot)
    base = str(tmp_path / "ck")
    opt2.run(6, isave=2, file_name=base)
    assert os.path.exists(base + "_odd.npz") and os.path.exists(base + "_even.npz")
    last = np.load(base + "_odd.npz")["coords"]
    assert last.shape == x0.shape
    opt2.reset(x0)
    assert opt2.get_niter() == 0
    with pytest.raises(ValueError):
        opt2.reset(np.zeros(5))


def test_combined_potential_sums(mods):
    _, _, NUwU, CombinedPotential = mods
    radii, Sk, V, x = nuwu_case(9)
    a = NUwU(radii, 0, Sk, V, np.ones(2), 1e-12)
    b = NUwU(radii, 0, Sk + 0.5, np.ones_like(V), np.ones(2), 1e-12)
    c = CombinedPotential([a, b])
    Ea, ga = a.get_energy_gradient(x)
    Eb, gb = b.get_energy_gradient(x)
    Ec, gc = c.get_energy_gradient(x)
    assert np.isclose(Ec, Ea + Eb, rtol=1e-14)
    assert np.allclose(gc, ga + gb, rtol=1e-13, atol=0)
    assert np.isclose(c.get_energy(x), Ea + Eb, rtol=1e-12)


def test_2d_n4096_lbfgs_teacher_forced(mods):
    """2D NUwU, N=4096, 65^2 modes, stealthy 0<|k|<=16, eps 1e-9, LBFGS M=10:
    every 50 iterations the device energy / gradient at the device iterate
    equal the oracle's exact direct sum at that iterate."""
    LBFGS_CPP, _, NUwU, _ = mods
    rng = np.random.default_rng(0)
    N, kp = 4096, 32
    k = D.mode_grid(kp, 2)
    kk = np.linalg.norm(k, axis=1)
    V = ((kk > 0) & (kk <= 16)).astype(float)
    Sk = np.zeros_like(V)
    x0 = rng.uniform(0, 1, 2 * N)
    pot = NUwU(np.ones(N), kp, Sk, V, np.ones(2), 1e-9, precision="fp64")
    ref = D.NUwU(np.ones(N), kp, Sk, V, np.ones(2))
    opt = LBFGS_CPP(x0, pot, tol=1e-8 / np.sqrt(2 * N), M=10, nsteps=2000)
    E0, g0 = ref.get_energy_gradient(x0)
    # NUFFT errors are relative to the input scale, not to the vanishing
    # energy near the stealthy optimum: compare against the initial scale
    while True:
        res = opt.run(20)
        E, g = ref.get_energy_gradient(res.coords)
        print("2D N=4096 it", res.nsteps, "E", res.energy, "oracle", E, "rms", res.rms, flush=True)
        assert abs(res.energy - E) <= 1e-8 * abs(E0)
        assert np.linalg.norm(res.grad - g) <= 1e-8 * np.linalg.norm(g0)
        if res.success or res.nsteps >= 2000:
            break
    assert res.success
    assert np.linalg.norm(g) < 1e-7


def test_2d_n4096_lbfgs_trajectory_vs_oracle(mods):
    """2D N = 4096, 65^2 modes, stealthy 0 < |k| <= 16, LBFGS M = 10: the
    device FP64 trajectory (eps 1e-12) against the oracle LBFGS
    (lbfgs.cpp restated) over the exact direct sums, iterate by iterate for
    100 iterations.

    The stealthy optimum is degenerate (a manifold of patterns with S = 0),
    so the LBFGS map amplifies any perturbation along it: a change of x0 at
    the 1e-12 level grows to ~1e-5 by iteration 60 in exact arithmetic.
    Positions therefore agree to 1e-8 for the first 30 iterations; after
    that the device's distance from the oracle trajectory must stay within
    10x the distance of a second oracle run started from x0 perturbed at the
    NUFFT tolerance (1e-12), i.e. the device behaves like an exact evaluation
    at that tolerance.  The line-search decisions (nfev) agree at every step
    and both runs end at the same S(k) on the constrained modes."""
    LBFGS_CPP, _, NUwU, _ = mods
    rng = np.random.default_rng(0)
    N, kp = 4096, 32
    k = D.mode_grid(kp, 2)
    kk = np.linalg.norm(k, axis=1)
    V = ((kk > 0) & (kk <= 16)).astype(float)
    Sk = np.zeros_like(V)
    x0 = rng.uniform(0, 1, 2 * N)
    pot = NUwU(np.ones(N), kp, Sk, V, np.ones(2), 1e-12, precision="fp64")
    ref = D.NUwU(np.ones(N), kp, Sk, V, np.ones(2))
    opt = LBFGS_CPP(x0, pot, tol=0.0, M=10)
    ora = OP.LBFGS(x0, ref, tol=0.0, M=10)
    ora.initialize_func_gradient()
    xp = x0 + 1e-12 * np.random.default_rng(1).standard_normal(x0.size)
    orp = OP.LBFGS(xp, ref, tol=0.0, M=10)
    orp.initialize_func_gradient()
    E0 = ora.f
    worst_early = 0.0
    for it in range(1, 101):
        r = opt.one_iteration()
        ora.one_iteration()
        orp.one_iteration()
        dev = float(np.max(np.abs(r.coords - ora.x)))
        env = float(np.max(np.abs(orp.x - ora.x)))
        if it % 10 == 0 or it <= 3:
            print("2D N=4096 it", it, "E", r.energy, "oracle", ora.f, "max|dx| device %.2e perturbed oracle %.2e"
                  % (dev, env), flush=True)
        assert r.nfev == ora.nfev and r.nsteps == ora.iter == it
        assert abs(r.energy - ora.f) <= 1e-10 * E0 + 1e-6 * abs(ora.f)
        if it <= 30:
            worst_early = max(worst_early, dev)
            assert dev <= 1e-8, (it, dev)
        else:
            assert dev <= 1e-8 + 10 * env, (it, dev, env)
    print("2D N=4096 first 30 iterates: max |dx| %.2e" % worst_early, flush=True)
    s_dev = np.abs(D.type1_grid(2 * np.pi * r.coords.reshape(-1, 2), np.ones(N), kp, 1)) ** 2 / N
    s_ora = np.abs(D.type1_grid(2 * np.pi * ora.x.reshape(-1, 2), np.ones(N), kp, 1)) ** 2 / N
    act = V > 0
    gap = float(np.max(np.abs(s_dev[act] - s_ora[act])))
    print("2D N=4096 final S(k) on the constrained modes: max %.2e / %.2e, gap %.2e"
          % (np.max(s_dev[act]), np.max(s_ora[act]), gap), flush=True)
    assert gap <= 1e-9 and r.energy < 1e-15 * E0


def test_checkpoints_alternate_and_events_track_oracle(mods, tmp_path):
    """run(niter, isave) (opt.pyx:41-54): after every `isave` iterations the
    coordinates go to <file>_odd, then <file>_even, alternately (the loop
    runs whole isave blocks, so it may pass niter); each file holds the
    iterate of its iteration exactly.  Events (opt.pyx:26-31, 56-65) see every
    iterate: their energies and coordinates equal the oracle LBFGS's at that
    iteration."""
    LBFGS_CPP, _, NUwU, _ = mods
    radii, Sk, V, x0 = nuwu_case(7)
    pot = NUwU(radii, 0, Sk, V, np.ones(2), 1e-13, precision="fp64")
    # the iterates of a plain one-iteration-at-a-time run
    walk = LBFGS_CPP(x0, pot, tol=0.0)
    iterates = {0: x0.copy()}
    for it in range(1, 13):
        iterates[it] = walk.one_iteration().coords.copy()
    for niter, isave, odd_it, even_it in ((3, 3, 3, None), (6, 3, 3, 6), (9, 3, 9, 6), (10, 3, 9, 12),
                                          (5, 2, 6, 4)):
        base = str(tmp_path / ("ck_%d_%d" % (niter, isave)))
        res = LBFGS_CPP(x0, pot, tol=0.0).run(niter, isave=isave, file_name=base)
        odd = np.load(base + "_odd.npz")["coords"]
        assert np.array_equal(odd, iterates[odd_it]), (niter, isave)
        if even_it is None:
            assert not os.path.exists(base + "_even.npz")
        else:
            assert np.array_equal(np.load(base + "_even.npz")["coords"], iterates[even_it]), (niter, isave)
        assert res.nsteps == max(odd_it, even_it or 0)
        assert np.array_equal(res.coords, iterates[res.nsteps])
    # events against the oracle LBFGS over the exact sums
    ref = D.NUwU(radii, 0, Sk, V, np.ones(2))
    ora = OP.LBFGS(x0, ref, tol=0.0, M=4)
    ora.initialize_func_gradient()
    seen = []
    opt = LBFGS_CPP(x0, pot, tol=0.0, events=[lambda coords, energy, rms: seen.append((coords.copy(), energy, rms))])
    opt.run(15)
    assert len(seen) == 15
    for it, (coords, energy, rms) in enumerate(seen, 1):
        ora.one_iteration()
        assert abs(energy - ora.f) <= 1e-10 * abs(ora.f), it
        assert np.allclose(coords, ora.x, rtol=0, atol=1e-10), it
        assert abs(rms - ora.rms) <= 1e-9 * ora.rms, it


def test_2d_nuwnu_n2p16_type3_lbfgs(mods):
    """NUwNU 2D, N = 2^16 points, 2^16 k uniform in |k| <= 128 (S0 = 0, V = 1)
    plus 32 peaks |k| in [20, 100] (S0 = N, V = 1): FP32 type-3 NUFFT against
    the exact FP64 sums (energy / gradient 1e-5 at the start and after LBFGS),
    and 500 LBFGS iterations (M = 10) lowering the energy."""
    from fresco_amd.potentials import NUwNU

    LBFGS_CPP = mods[0]
    rng = np.random.default_rng(0)
    N = M = 1 << 16
    r = 128 * np.sqrt(rng.uniform(0, 1, M))
    th = rng.uniform(0, 2 * np.pi, M)
    rp = rng.uniform(20, 100, 32)
    tp = rng.uniform(0, 2 * np.pi, 32)
    K = np.concatenate([np.stack([r * np.cos(th), r * np.sin(th)], 1), np.stack([rp * np.cos(tp), rp * np.sin(tp)], 1)])
    Sk = np.concatenate([np.zeros(M), np.full(32, float(N))])
    V = np.ones(M + 32)
    x0 = rng.uniform(0, 1, 2 * N)
    p32 = NUwNU(np.ones(N), K.ravel(), Sk, V, np.ones(2), 1e-6, precision="fp32")
    exact = NUwNU(np.ones(N), K.ravel(), Sk, V, np.ones(2), method="direct")

    def compare(x, scale_e, scale_g):
        E, g = p32.get_energy_gradient(x)
        Ee, ge = exact.get_energy_gradient(x)
        assert abs(E - Ee) <= 1e-5 * scale_e
        assert np.linalg.norm(g - ge) <= 1e-5 * scale_g
        return Ee, ge

    E0, g0 = exact.get_energy_gradient(x0)
    compare(x0, abs(E0), np.linalg.norm(g0))
    opt = LBFGS_CPP(x0, p32, M=10, nsteps=500, tol=0.0)
    res = opt.run(500)
    print("NUwNU 2D N=2^16 LBFGS", res.nsteps, "E", E0, "->", res.energy, flush=True)
    assert res.nsteps == 500
    assert res.energy < 0.8 * E0  # measured: 1.80e9 -> 1.21e9
    compare(res.coords, abs(E0), np.linalg.norm(g0))


@pytest.fixture(scope="module")
def c2(mods):
    """3D NUwU, N = 2^20, 129^3 modes, stealthy 0 < |k| <= 24, S0 = 0, V = 1:
    the FP64 plain potential (eps 1e-9), the FP32 adjoint one (eps 1e-6) and
    x0; the FP64-driven run's final coordinates are cached by the test that
    makes them."""
    _, _, NUwU, _ = mods
    N, kp = 1 << 20, 64
    k = D.mode_grid(kp, 3)
    kk = np.linalg.norm(k, axis=1)
    V = ((kk > 0) & (kk <= 24)).astype(float)
    Sk = np.zeros_like(V)
    x0 = np.random.default_rng(0).uniform(0, 1, 3 * N)
    p64 = NUwU(np.ones(N), kp, Sk, V, np.ones(3), 1e-9, precision="fp64")
    p32 = NUwU(np.ones(N), kp, Sk, V, np.ones(3), 1e-6, precision="fp32")
    return {"N": N, "x0": x0, "p64": p64, "p32": p32, "act": V > 0, "final64": None}


def _c2_eval(c2, xh):
    """FP64 and FP32 energy, gradient and S(k) on the active modes at xh."""
    import torch

    N, act = c2["N"], c2["act"]
    x = torch.tensor(xh, device="cuda")
    g64 = torch.empty_like(x)
    g32 = torch.empty_like(x)
    e64 = c2["p64"].energy_gradient_device(x, g64)
    e32 = c2["p32"].energy_gradient_device(x, g32)
    s64 = (np.abs(c2["p64"].get_rho()) ** 2 / N)[act]
    s32 = (np.abs(c2["p32"].get_rho()) ** 2 / N)[act]
    return e64, e32, g64, g32, s64, s32


def test_3d_n2p20_lbfgs_teacher_forced(mods, c2):
    """3D N = 2^20 configuration: 200 LBFGS iterations (M = 4) on the FP64
    plain path; at every one of the 200 iterates the FP32 adjoint path is
    evaluated at the FP64 iterate (teacher forcing).  The NUFFT's error is
    relative to the input scale (sum |c| = N), not to the energy, gradient and
    S(k) that vanish towards the stealthy optimum, so the errors are measured
    against their values at x0 (within 1e-5 at every iterate), and each
    relative error is held to 1e-5 while its quantity is above 0.3 of its
    value at x0 (the error is ~2.6e-6 of the x0 scale)."""
    import torch

    LBFGS_CPP = mods[0]
    x0 = c2["x0"]
    e0, _, g0, _, s0, _ = _c2_eval(c2, x0)
    gn0 = float(torch.linalg.norm(g0))
    sn0 = float(np.linalg.norm(s0))
    opt = LBFGS_CPP(x0, c2["p64"], M=4, tol=0.0, nsteps=200)
    worst = [0.0, 0.0, 0.0]
    for it in range(200):
        res = opt.one_iteration()
        e64, e32, g64, g32, s64, s32 = _c2_eval(c2, res.coords)
        assert abs(e64 - res.energy) <= 1e-12 * abs(e64)  # the optimizer's energy is the FP64 path's
        de = abs(e32 - e64)
        dg = float(torch.linalg.norm(g32 - g64))
        ds = float(np.linalg.norm(s32 - s64))
        scaled = (de / e0, dg / gn0, ds / sn0)
        worst = [max(a, b) for a, b in zip(worst, scaled)]
        rel = (de / abs(e64), dg / float(torch.linalg.norm(g64)), ds / np.linalg.norm(s64))
        if it % 25 == 24 or it < 3:
            print("3D N=2^20 it", res.nsteps, "E", e64, "rel E, g, S", rel, "vs x0 scale", scaled, flush=True)
        assert max(scaled) <= 1e-5, (it, scaled)
        for name, r_, now, start in (("E", rel[0], abs(e64), e0), ("g", rel[1], float(torch.linalg.norm(g64)), gn0),
                                     ("S", rel[2], float(np.linalg.norm(s64)), sn0)):
            if now > 0.3 * start:
                assert r_ < 1e-5, (it, name, r_)
    print("3D N=2^20 worst error over 200 iterates vs x0 scale (E, g, S):", worst, flush=True)
    assert res.nsteps == 200 and res.energy < 1e-6 * e0
    c2["final64"] = res.coords.copy()
    c2["e0"] = e0


def test_3d_n2p20_fp32_driven_lbfgs_converged_sk(mods, c2):
    """The same 200 LBFGS iterations (M = 4, same x0) driven by the FP32
    adjoint path itself.  Both runs' final patterns are evaluated on the FP64
    plain path: the FP32-driven run must reach the FP64-driven run's
    converged S(k) on every constrained mode to 1e-6 (absolute, S(k) is 0 at
    the target) and an energy below 1e-6 of its start."""
    LBFGS_CPP = mods[0]
    if c2["final64"] is None:
        opt64 = LBFGS_CPP(c2["x0"], c2["p64"], M=4, tol=0.0, nsteps=200)
        c2["final64"] = opt64.run(200).coords.copy()
        c2["e0"] = c2["p64"].get_energy(c2["x0"])
    e0 = c2["e0"]
    opt = LBFGS_CPP(c2["x0"], c2["p32"], M=4, tol=0.0, nsteps=200)
    for block in range(8):
        res = opt.run(25)
        e64 = c2["p64"].get_energy(res.coords)
        print("FP32-driven it", res.nsteps, "E fp32", res.energy, "E fp64 at the iterate", e64, flush=True)
    assert res.nsteps == 200
    e32run, _, _, _, s_a, _ = _c2_eval(c2, res.coords)
    e64run, _, _, _, s_b, _ = _c2_eval(c2, c2["final64"])
    gap = float(np.max(np.abs(s_a - s_b)))
    print("converged S(k) on the constrained modes: FP32-driven max %.3e, FP64-driven max %.3e, max |gap| %.3e; "
          "FP64 energy of the final patterns: FP32-driven %.3e, FP64-driven %.3e (start %.3e)"
          % (np.max(s_a), np.max(s_b), gap, e32run, e64run, e0), flush=True)
    assert gap <= 1e-6
    assert e32run < 1e-6 * e0
