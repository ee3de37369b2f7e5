"""The reference's example notebooks (examples/NUwU.ipynb, NUwNU.ipynb,
UwU.ipynb, UwNU.ipynb), cell by cell with their own imports
(``from fresco.potentials import ...``, ``from fresco.optimize import
LBFGS_CPP``, ``from fresco.utils import export_structure``), at reduced sizes
and with seeded inputs (the notebooks draw unseeded ones), on the GPU.  The
results are checked against the oracle's exact sums at the final patterns."""
import numpy as np
import pytest

from oracle import sf_direct as D
from oracle import structure as ST

pytestmark = pytest.mark.gpu


def test_nuwu_notebook_sequence():
    from fresco.optimize import LBFGS_CPP
    from fresco.potentials import NUwU
    from fresco.utils import export_structure

    # cell 1 (N = 10000 in the notebook)
    N = 2000
    ndim = 2
    L = np.ones(ndim)
    c = np.ones(N)
    x0 = np.random.default_rng(0).random(N * ndim)
    # cell 2 (K = 80 in the notebook)
    K = 20
    Kvec = np.arange(-K, K + 1).reshape((1, -1))
    Kmag = np.sqrt(Kvec ** 2 + (Kvec ** 2).T)
    Kmask = np.array(Kmag <= K, dtype=np.int32)
    # cell 3
    Sk = np.zeros(Kmag.shape)
    wk = Kmask
    center = int(Kmask.shape[0] / 2)
    wk[center, center] = 0
    # cell 4
    loss = NUwU(c, K, Sk, wk, L)
    minimizer = LBFGS_CPP(x0, loss, iprint=100)
    # cell 5 (run(1000) in the notebook)
    minimizer.run(300)
    res = minimizer.get_result()
    x = np.array(res.coords)
    # cell 6
    points = x.reshape(-1, ndim)
    # cell 7
    Kspan = K * 10
    q, S, Skg, r, g, gr = export_structure(points, L, Nk=Kspan)

    ref = D.NUwU(c, K, Sk, wk, L)
    E0 = ref.get_energy(x0)
    E, grad = ref.get_energy_gradient(x)
    print("NUwU notebook: E", E0, "->", res.energy, "(exact", E, ") rms", res.rms, flush=True)
    # NUwU's default eps = 1e-6 (nuwu.pyx:11): errors at that level of the x0 scale
    assert abs(res.energy - E) <= 1e-6 * E0
    assert np.linalg.norm(np.asarray(res.grad) - grad) <= 1e-6 * np.linalg.norm(ref.get_energy_gradient(x0)[1])
    assert E < 1e-8 * E0
    # the final S(k) on the constrained disk, from export_structure's grid
    # (first index along x, k = index - Kspan // 2) and from the exact sums
    kx, ky = np.meshgrid(np.arange(Kspan) - Kspan // 2, np.arange(Kspan) - Kspan // 2, indexing="ij")
    disk = (np.hypot(kx, ky) <= K) & (np.hypot(kx, ky) > 0)
    assert Skg.shape == (Kspan, Kspan)
    # the loss is the sum of S(k)^2 over the constrained disk (S0 = 0, V = 1)
    assert abs(np.sum(Skg[disk] ** 2) - E) <= 1e-4 * E
    S_exact = ST.structure_factor(points, L, Kspan)
    # export_structure's tol = 1e-9 is relative to the input scale (sum c = N)
    print("NUwU notebook: export_structure max |S - exact| %.2e" % np.max(np.abs(Skg - S_exact)), flush=True)
    assert np.max(np.abs(Skg - S_exact)) <= 20 * 1e-9 * N
    assert q.shape == S.shape and r.shape == g.shape


def test_nuwnu_notebook_sequence():
    from fresco.optimize import LBFGS_CPP
    from fresco.potentials import NUwNU
    from fresco.utils import export_structure

    N = 2000
    ndim = 2
    L = np.ones(ndim)
    c = np.ones(N)
    x0 = np.random.default_rng(1).random(N * ndim)
    K = 100
    Npeaks = 5
    thetas = (np.arange(Npeaks) / Npeaks) * np.pi
    Kvecs = np.vstack([np.cos(thetas), np.sin(thetas)]).T * K
    Sk = np.ones(Kvecs.shape[0]) * N
    wk = np.ones(Kvecs.shape[0])
    loss = NUwNU(c, Kvecs.ravel(), Sk, wk, L)
    minimizer = LBFGS_CPP(x0.ravel(), loss, iprint=100)
    minimizer.run(200)
    res = minimizer.get_result()
    x = np.array(res.coords)
    points = x.reshape(-1, ndim)
    hamming = 0.54 + 0.46 * np.cos(np.linalg.norm(points, axis=-1) * 2 * np.pi / np.sqrt(ndim))
    Kspan = 400  # K * 10 in the notebook
    q, S, Skg, r, g, gr = export_structure(points, L, Nk=Kspan, radii=hamming)

    ref = D.NUwNU(c, Kvecs.ravel(), Sk, wk, L)
    E0 = ref.get_energy(x0)
    E = ref.get_energy(x)
    print("NUwNU notebook: E", E0, "->", res.energy, "(exact", E, ")", flush=True)
    assert abs(res.energy - E) <= 1e-6 * E0  # default eps = 1e-6 (nuwnu.pyx:11)
    assert E < 0.9 * E0
    err = np.max(np.abs(Skg - ST.structure_factor(points, L, Kspan, radii=hamming)))
    print("NUwNU notebook: export_structure max |S - exact| %.2e" % err, flush=True)
    assert err <= 20 * 1e-9 * N


def test_uwu_notebook_sequence():
    from fresco.optimize import LBFGS_CPP
    from fresco.potentials import UwU

    Nside = 20  # 100 in the notebook
    ndim = 2
    L = np.ones(ndim)
    gridsize = np.ones(ndim, dtype=np.int32) * int(2 * Nside + 1)
    x0 = np.random.default_rng(2).random(gridsize)
    K = 8
    Kvec = np.arange(-Nside, Nside + 1).reshape((1, -1))
    Kmag = np.sqrt(Kvec ** 2 + (Kvec ** 2).T)
    Kmask = np.array(Kmag <= K, dtype=np.int32) * np.array(Kmag >= K * 0.9, dtype=np.int32)
    Sk = np.ones(Kmag.shape) * 10
    wk = Kmask
    center = int(Kmask.shape[0] / 2)
    wk[center, center] = 0
    loss = UwU(gridsize, K, Sk, wk, L)
    minimizer = LBFGS_CPP(x0.ravel(), loss, iprint=100)
    minimizer.run(300)
    res = minimizer.get_result()
    x = np.array(res.coords)
    output = x.reshape(gridsize)
    rhohat = np.fft.fftshift(np.fft.fft2(output))
    Skn = np.absolute(rhohat) ** 2
    Skn /= np.sum(np.absolute(output) ** 2)

    ref = D.UwU(gridsize, K, Sk, wk, L)
    E0 = ref.get_energy(x0.ravel())
    E, grad = ref.get_energy_gradient(x)
    print("UwU notebook: E", E0, "->", res.energy, "(exact", E, ")", flush=True)
    assert abs(res.energy - E) <= 1e-10 * max(E0, E)
    assert np.allclose(res.grad, grad, rtol=0, atol=1e-10 * np.max(np.abs(grad)) + 1e-300)
    assert np.all(np.isfinite(Skn))


def test_uwnu_notebook_sequence():
    from fresco.optimize import LBFGS_CPP
    from fresco.potentials import UwNU

    Nside = 20  # 100 in the notebook
    ndim = 2
    L = np.ones(ndim)
    gridsize = np.ones(ndim, dtype=np.int32) * int(2 * Nside + 1)
    x0 = np.ones(gridsize)
    K = 4
    Npeaks = 5
    thetas = (np.arange(Npeaks) / Npeaks) * np.pi
    Kvecs = np.vstack([np.cos(thetas), np.sin(thetas)]).T * K / Nside
    Kvecs = np.vstack([np.zeros(ndim), Kvecs])
    Sk = np.ones(Kvecs.shape[0]) * np.sum(x0)
    wk = np.ones(Kvecs.shape[0])
    loss = UwNU(gridsize, Kvecs, Sk, wk, L)
    minimizer = LBFGS_CPP(x0.ravel(), loss, iprint=100)
    minimizer.run(300)
    res = minimizer.get_result()
    x = np.array(res.coords)

    ref = D.UwNU(gridsize, Kvecs, Sk, wk, L)
    E0 = ref.get_energy(x0.ravel())
    E, grad = ref.get_energy_gradient(x)
    print("UwNU notebook: E", E0, "->", res.energy, "(exact", E, ")", flush=True)
    # default eps = 1e-6 (uwnu.pyx:11): errors at that level of the x0 scale
    assert abs(res.energy - E) <= 1e-6 * max(E0, E)
    assert np.linalg.norm(np.asarray(res.grad) - grad) <= 1e-6 * np.linalg.norm(ref.get_energy_gradient(x0.ravel())[1])
    assert E < E0
