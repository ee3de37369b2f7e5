"""LBFGS iterations replayed from a recorded HIP graph against the same
kernels launched one by one.

Both optimizers drive ONE potential object and alternate: every replay is
followed by eager work on that potential (an evaluation at an unrelated x and
the launch-by-launch optimizer's own iteration, each ending in a host
synchronisation), and after every iteration the two trajectories must agree
bit for bit (coordinates, energy, H0, rms, evaluation count).  The recorded
path is the product default; the launch-by-launch path is the reference
structure of lbfgs.cpp:31-64 (its parity with the optimizer oracle is covered
in test_optim_gpu.py)."""
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


def make_case(P, name):
    rng = np.random.default_rng(7)
    if name == "nuwu2d_fp64":
        N, kp = 400, 12
        Sk, V = stealthy(kp, 2, 6)
        pot = P.NUwU(np.ones(N), kp, Sk, V, np.ones(2), 1e-12, precision="fp64")
        return pot, rng.uniform(0, 1, 2 * N)
    if name == "nuwu2d_fp32":
        N, kp = 4096, 32
        Sk, V = stealthy(kp, 2, 16)
        pot = P.NUwU(np.ones(N), kp, Sk, V, np.ones(2), 1e-6, precision="fp32")
        return pot, rng.uniform(0, 1, 2 * N)
    if name == "nuwu3d_fp32":
        N, kp = 3000, 10
        Sk, V = stealthy(kp, 3, 6)
        pot = P.NUwU(np.ones(N), kp, Sk, V, np.ones(3), 1e-6, precision="fp32")
        return pot, rng.uniform(0, 1, 3 * N)
    if name == "batch2d_fp32":
        pats, xs = [], []
        for N, n in ((60, 9), (300, 17), (150, 13)):
            Sk, V = stealthy(n // 2, 2, n // 3)
            pats.append((np.ones(N), Sk, V))
            xs.append(rng.uniform(0, 1, 2 * N))
        return P.NUwUBatch(pats, np.ones(2), 1e-6, precision="fp32"), np.concatenate(xs)
    if name.startswith("uwu"):
        shape, prec = ((24, 20), "fp64") if name == "uwu2d_fp64" else ((12, 10, 8), "fp32")
        Sk = rng.uniform(0.5, 3, shape)
        Sk.flat[::4] = 0.0
        V = rng.uniform(0, 1, shape)
        pot = P.UwU(np.array(shape), 0, Sk, V, np.ones(len(shape)), precision=prec)
        return pot, rng.uniform(0.2, 1.8, int(np.prod(shape)))
    if name == "uwnu2d_fp32":
        shape, nk = (16, 12), 150
        K = rng.uniform(-1.3, 1.3, (nk, 2))
        Sk = rng.uniform(0.5, 3, nk)
        V = rng.uniform(0, 1, nk)
        pot = P.UwNU(np.array(shape), K, Sk, V, np.ones(2), 1e-6, precision="fp32")
        return pot, rng.uniform(0.2, 1.8, int(np.prod(shape)))
    if name.startswith("nuwnu"):
        N, nk = 500, 300
        K = rng.uniform(-20, 20, (nk, 2))
        Sk = rng.uniform(0, 2, nk)
        V = rng.uniform(0, 1, nk)
        method = "direct" if name == "nuwnu2d_direct" else "nufft"
        pot = P.NUwNU(rng.uniform(0.5, 1.5, N), K.ravel(), Sk, V, np.ones(2), 1e-6, precision="fp32",
                      method=method)
        return pot, rng.uniform(0, 1, 2 * N)
    if name in ("pair3d_cells", "wca2d_allpairs", "combined2d"):
        d, N = (3, 2000) if name == "pair3d_cells" else (2, 600)
        radii = rng.uniform(0.45, 0.55, N)
        vol = np.sum(np.pi * radii ** 2) if d == 2 else np.sum(4 / 3 * np.pi * radii ** 3)
        box = np.full(d, (vol / 0.6) ** (1.0 / d))
        x0 = rng.uniform(0, 1, (N, d)) * box
        if name == "wca2d_allpairs":
            return P.WCAPotential(1.0, 1.0, radii, box), x0.ravel()
        pair = P.InversePower(2.5, 1.0, radii, box, use_cell_lists=True)
        if name == "pair3d_cells":
            return pair, x0.ravel()
        Sk, V = stealthy(8, 2, 4)
        nuwu = P.NUwU(np.ones(N), 8, Sk, V, box, 1e-12, precision="fp64")
        return P.CombinedPotential([nuwu, pair]), x0.ravel()
    raise ValueError(name)


def same(ra, rb):
    assert np.array_equal(ra.coords, rb.coords)
    assert ra.energy == rb.energy and ra.rms == rb.rms
    assert ra["H0"] == rb["H0"] and ra.nfev == rb.nfev and ra.nsteps == rb.nsteps


@pytest.mark.parametrize("name", ["nuwu2d_fp64", "nuwu2d_fp32", "nuwu3d_fp32", "batch2d_fp32", "uwu2d_fp64",
                                  "uwu3d_fp32", "uwnu2d_fp32", "nuwnu2d_type3", "nuwnu2d_direct", "pair3d_cells",
                                  "wca2d_allpairs", "combined2d"])
def test_graph_replays_match_launches(mods, name):
    LBFGS_CPP, P = mods
    pot, x0 = make_case(P, name)
    print("case", name, "ndof", len(x0), flush=True)
    ga = LBFGS_CPP(x0, pot, tol=0.0, M=4, use_graph=True)
    eb = LBFGS_CPP(x0, pot, tol=0.0, M=4, use_graph=False)
    xr = np.random.default_rng(1).uniform(0, 1, len(x0))
    for it in range(30):
        ra = ga.one_iteration()
        pot.get_energy_gradient(xr)  # eager work between replays
        rb = eb.one_iteration()
        same(ra, rb)
    assert ga.graph_stats() == (1, 30)
    assert eb.graph_stats() == (0, 0)
    print(name, "E", ra.energy, "rms", ra.rms, "nfev", ra.nfev, flush=True)


def test_graph_backtracking_and_memory_wrap(mods):
    """Demanding a 60% relative energy drop per step (a negative maxErise with
    rel_energy) makes first trials fail: the host continues the backtracking
    after the replay, sometimes through all 10 trials; M=3 wraps the memory
    ring many times."""
    LBFGS_CPP, P = mods
    pot, x0 = make_case(P, "nuwu2d_fp64")
    kw = dict(tol=0.0, M=3, maxstep=1.0, maxErise=-0.6, rel_energy=True)
    ga = LBFGS_CPP(x0, pot, use_graph=True, **kw)
    eb = LBFGS_CPP(x0, pot, use_graph=False, **kw)
    for it in range(40):
        ra = ga.one_iteration()
        rb = eb.one_iteration()
        same(ra, rb)
    print("nfev", ra.nfev, "iterations", ra.nsteps, flush=True)
    assert ra.nfev > ra.nsteps + 20  # backtracking happened (oracle: 140 / 40)
    assert ga.graph_stats() == (1, 40)


def test_graph_rerecorded_when_targets_grow(mods):
    """set_targets switching on modes beyond the active band rebuilds the fine
    grid: the recorded iteration is stale and must be recorded again."""
    LBFGS_CPP, P = mods
    N, kp = 500, 12
    Sk, V = stealthy(kp, 2, 4)
    pot = P.NUwU(np.ones(N), kp, Sk, V, np.ones(2), 1e-6, precision="fp32")
    x0 = np.random.default_rng(3).uniform(0, 1, 2 * N)
    ga = LBFGS_CPP(x0, pot, tol=0.0, M=4, use_graph=True)
    eb = LBFGS_CPP(x0, pot, tol=0.0, M=4, use_graph=False)
    for it in range(8):
        same(ga.one_iteration(), eb.one_iteration())
    nf0 = pot.plan_info()["nf"]
    Sk2, V2 = stealthy(kp, 2, 10)
    pot.set_targets(Sk2, V2)
    assert pot.plan_info()["nf"] != nf0
    for it in range(8):
        same(ga.one_iteration(), eb.one_iteration())
    assert ga.graph_stats() == (2, 16)


def test_graph_off_for_unrecordable_potentials(mods):
    """Potentials without an enqueue-only evaluation (here a Python one, alone
    and inside a combined potential) run the launch-by-launch path with the
    same results."""
    LBFGS_CPP, P = mods

    class Quartic:
        def get_energy_gradient(self, x):
            x = np.asarray(x)
            return float(np.sum(x ** 4)), 4 * x ** 3

        def get_energy(self, x):
            return float(np.sum(np.asarray(x) ** 4))

    x0 = np.random.default_rng(5).uniform(-1, 1, 50)
    ga = LBFGS_CPP(x0, Quartic(), tol=0.0, M=4, use_graph=True)
    eb = LBFGS_CPP(x0, Quartic(), tol=0.0, M=4, use_graph=False)
    for it in range(20):
        same(ga.one_iteration(), eb.one_iteration())
    assert ga.graph_stats() == (0, 0)
    # a combined potential with a host term is not recorded either
    box = np.full(2, 6.0)
    xc = np.random.default_rng(6).uniform(0, 6, 50)
    comb = P.CombinedPotential([P.InversePower(2.5, 1.0, np.full(25, 0.5), box, use_cell_lists=True),
                                P.PythonPotential(Quartic(), 50)])
    gc = LBFGS_CPP(xc, comb, tol=0.0, M=4, use_graph=True)
    ec = LBFGS_CPP(xc, comb, tol=0.0, M=4, use_graph=False)
    for it in range(10):
        same(gc.one_iteration(), ec.one_iteration())
    assert gc.graph_stats() == (0, 0)


def _fire_same(ra, rb):
    assert np.array_equal(ra.coords, rb.coords)
    assert ra.energy == rb.energy and ra.rms == rb.rms
    assert ra.nfev == rb.nfev and ra.nsteps == rb.nsteps


@pytest.mark.parametrize("name,stepback", [("nuwu2d_fp64", True), ("nuwu2d_fp32", True), ("nuwu3d_fp32", False),
                                           ("nuwu3d_fp32", True), ("uwu3d_fp32", True), ("batch2d_fp32", True),
                                           ("pair3d_cells", True), ("combined2d", True)])
def test_fire_graph_replays_match_host_iterations(mods, name, stepback):
    """Modified FIRE with its scalars on the device, each iteration replayed
    from one of two recorded graphs (the position / gradient buffer pairs
    swap roles every accepted step), against the host-controlled iteration
    of modified_fire.hpp:136-216 on the same potential: bitwise equal
    trajectories.  The large time step makes some steps go uphill (power
    P <= 0: velocity reset, time step cut, and with stepback the previous
    position restored)."""
    from fresco_amd.optimize import ModifiedFireCPP

    _, P = mods
    pot, x0 = make_case(P, name)
    kw = dict(dtstart=0.02, dtmax=0.2, maxstep=0.05, Nmin=2, tol=0.0, stepback=stepback)
    ga = ModifiedFireCPP(x0, pot, use_graph=True, **kw)
    eb = ModifiedFireCPP(x0, pot, use_graph=False, **kw)
    xr = np.random.default_rng(2).uniform(0, 1, len(x0))
    energies = []
    for it in range(40):
        ra = ga.one_iteration()
        pot.get_energy_gradient(xr)  # eager work between replays
        rb = eb.one_iteration()
        _fire_same(ra, rb)
        energies.append(ra.energy)
    cap, rep = ga.graph_stats()
    assert rep == 40 and 1 <= cap <= 2
    assert eb.graph_stats() == (0, 0)
    rises = int(np.sum(np.diff(energies) > 0))
    backs = int(np.sum(np.diff(energies) == 0))  # stepback restores the previous energy exactly
    print(name, "stepback", stepback, "graphs", cap, "energy rises", rises, "steps back", backs, "E", energies[-1],
          flush=True)
    if name == "nuwu3d_fp32":
        # this case has uphill steps: the recorded branch is exercised
        assert (backs if stepback else rises) >= 1
