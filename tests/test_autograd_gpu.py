"""torch.autograd through the structure-factor potentials.

* small cases: autograd gradients vs the oracle's exact gradient and
  target derivatives (relative l2 1e-9 in FP64, 1e-5 in FP32);
* differentiable S(k) shaping, 2D, N = 2^18, 257^2 modes, S0 = sum of 8
  Gaussians: FP32 forward + backward against central differences of the
  FP64 path at 64 random coordinates (h = 1e-4, and its Richardson
  extrapolation) and at 64 random modes of S0;
* UwU (exact_grad) and UwNU coordinate gradients vs finite differences.
"""
import math

import numpy as np
import pytest
import torch

from oracle import sf_direct as O

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi


def rel(a, b):
    a = np.ravel(np.asarray(a))
    b = np.ravel(np.asarray(b))
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def mods():
    import fresco_amd.autograd as A
    import fresco_amd.potentials as P

    return A, P


# V(k = 0) != 0 here: the adjoint paths keep the k = 0 constant off their
# derivative-interpolated grid (it moves no point)
@pytest.mark.parametrize("prec,tol,eps", [("fp64", 1e-9, 1e-12), ("fp64-adjoint", 1e-9, 1e-12), ("fp32", 1e-5, 1e-6)])
def test_nuwu_autograd_vs_oracle(mods, prec, tol, eps):
    A, P = mods
    rng = np.random.default_rng(21)
    d, n, N = 2, 17, 300
    M = n ** d
    radii = rng.uniform(0.6, 1.4, N)
    Sk = rng.uniform(0.2, 2, M)
    Sk[::3] = 0
    V = rng.uniform(0, 1, M)
    V[::7] = 0
    L = np.array([1.0, 1.5])
    p = rng.uniform(0, 1, N * d)
    pot = P.NUwU(radii, 0, np.ones(M), np.ones(M), L, eps, precision=prec)
    x = torch.tensor(p, device="cuda", requires_grad=True)
    S0 = torch.tensor(Sk, requires_grad=True)
    Vt = torch.tensor(V, requires_grad=True)
    E = A.structure_factor_loss(pot, x, S0, Vt)
    E.backward()
    ref = O.NUwU(radii, 0, Sk, V, L)
    E0, g0 = ref.get_energy_gradient(p)
    assert abs(E.item() - E0) / abs(E0) < tol
    assert rel(x.grad.cpu().numpy(), g0 * np.tile(TWO_PI / L, N)) < tol
    dS0, dV = O.target_grads(ref.rho(p), Sk, V, N)
    assert S0.grad.shape == S0.shape and S0.grad.device == S0.device
    assert rel(S0.grad.numpy(), dS0) < tol
    assert rel(Vt.grad.numpy(), dV) < tol
    # the targets stay set for plain calls
    assert abs(pot.get_energy(p) - E0) / abs(E0) < tol


def gaussians_s0(kp, rng, ng=8):
    k = O.mode_grid(kp, 2).astype(float)
    r = 100 * np.sqrt(rng.uniform(0, 1, ng))
    th = rng.uniform(0, TWO_PI, ng)
    c = np.stack([r * np.cos(th), r * np.sin(th)], 1)
    w = rng.uniform(3, 10, ng)
    a = rng.uniform(0, 2, ng)
    S0 = np.zeros(len(k))
    for g in range(ng):
        S0 += a[g] * np.exp(-np.sum((k - c[g]) ** 2, 1) / (2 * w[g] ** 2))
    # the relative branch divides by S0: Gaussian tails (down to 1e-300) would
    # overflow the loss, so S0 < 1e-3 takes the absolute branch (S0 = 0)
    S0[S0 < 1e-3] = 0.0
    V = np.ones(len(k))
    V[len(k) // 2] = 0.0  # k = 0
    return S0, V


def test_sk_shaping_2d_fd(mods):
    """N = 2^18, 257^2 modes, 8-Gaussian S0 (tails below 1e-3 set to 0),
    V = 1 for k != 0."""
    A, P = mods
    rng = np.random.default_rng(0)
    N, d, kp = 1 << 18, 2, 128
    S0np, Vnp = gaussians_s0(kp, rng)
    p = rng.uniform(0, 1, N * d)
    L = np.ones(d)
    p32 = P.NUwU(np.ones(N), kp, S0np, Vnp, L, 1e-6, precision="fp32")
    p64 = P.NUwU(np.ones(N), kp, S0np, Vnp, L, 1e-12, precision="fp64")
    x = torch.tensor(p, device="cuda", requires_grad=True)
    S0 = torch.tensor(S0np, device="cuda", requires_grad=True)
    E = A.structure_factor_loss(p32, x, S0, torch.tensor(Vnp, device="cuda"))
    E.backward()
    gx = x.grad.cpu().numpy()
    gS = S0.grad.cpu().numpy()

    E64, g64 = p64.get_energy_gradient(p)
    print("2D N=2^18 energies fp32/fp64", E.item(), E64, flush=True)
    assert abs(E.item() - E64) / E64 < 1e-5
    assert rel(gx, g64 * TWO_PI) < 1e-5

    idx = rng.choice(N * d, 64, replace=False)

    def dE(i, h):
        q = p.copy()
        q[i] += h
        ep = p64.get_energy(q)
        q[i] -= 2 * h
        return (ep - p64.get_energy(q)) / (2 * h)

    fd = np.array([dE(i, 1e-4) for i in idx])
    fd2 = np.array([dE(i, 5e-5) for i in idx])
    rich = (4 * fd2 - fd) / 3
    r1, r2 = rel(gx[idx], fd), rel(gx[idx], rich)
    print("2D N=2^18 x-grad vs FD h=1e-4:", r1, " Richardson:", r2, flush=True)
    assert r1 < 2e-3  # central difference truncation O((2 pi k h)^2)
    assert r2 < 1e-5

    # d phi / d S0 at 64 random modes with S0 != 0, central differences of FP64
    cand = np.flatnonzero((S0np > 1e-3) & (Vnp != 0))
    modes = rng.choice(cand, 64, replace=False)
    p64.get_energy(p)
    fdS = []
    for mdx in modes:
        h = 1e-4 * S0np[mdx]
        s = S0np.copy()
        s[mdx] += h
        p64.set_targets(s, Vnp)
        ep = p64.get_energy(p)
        s[mdx] -= 2 * h
        p64.set_targets(s, Vnp)
        em = p64.get_energy(p)
        fdS.append((ep - em) / (2 * h))
    p64.set_targets(S0np, Vnp)
    rS = rel(gS[modes], np.array(fdS))
    print("2D N=2^18 S0-grad vs FD:", rS, flush=True)
    assert rS < 1e-5


def test_uwu_exact_and_uwnu_autograd_fd(mods):
    A, P = mods
    rng = np.random.default_rng(5)
    shape = (6, 5)
    Sk = rng.uniform(0.5, 3, shape)
    V = rng.uniform(0, 1, shape)
    xg = rng.uniform(0.5, 1.5, 30)
    pot = P.UwU(np.array(shape), 0, Sk, V, np.ones(2), exact_grad=True)
    x = torch.tensor(xg, requires_grad=True)
    S0 = torch.tensor(Sk, requires_grad=True)
    E = A.structure_factor_loss(pot, x, S0, torch.tensor(V))
    E.backward()
    h = 1e-6
    fd = []
    for i in range(30):
        e = np.zeros(30)
        e[i] = h
        fd.append((pot.get_energy(xg + e) - pot.get_energy(xg - e)) / (2 * h))
    assert rel(x.grad.numpy(), fd) < 1e-6
    # S0 grads in the caller's centred layout
    fdS = np.zeros(shape)
    for i in np.ndindex(shape):
        s = Sk.copy()
        s[i] += h
        pot.set_targets(s, V)
        ep = pot.get_energy(xg)
        s[i] -= 2 * h
        pot.set_targets(s, V)
        fdS[i] = (ep - pot.get_energy(xg)) / (2 * h)
    assert rel(S0.grad.numpy(), fdS) < 1e-6
    with pytest.raises(ValueError):
        A.structure_factor_loss(P.UwU(np.array(shape), 0, Sk, V, np.ones(2)), torch.tensor(xg, requires_grad=True))

    shape = (7, 6)
    nk = 40
    K = rng.uniform(-2, 2, (nk, 2))
    Sk = rng.uniform(0.5, 2, nk)
    Vk = rng.uniform(0, 1, nk)
    xg = rng.uniform(0.5, 1.5, 42)
    pot = P.UwNU(np.array(shape), K, Sk, Vk, np.ones(2), 1e-12)
    x = torch.tensor(xg, requires_grad=True)
    E = A.structure_factor_loss(pot, x)
    E.backward()
    fd = []
    for i in range(42):
        e = np.zeros(42)
        e[i] = h
        fd.append((pot.get_energy(xg + e) - pot.get_energy(xg - e)) / (2 * h))
    assert rel(x.grad.numpy(), fd) < 1e-6


def test_nuwnu_target_grads(mods):
    A, P = mods
    rng = np.random.default_rng(6)
    N, nk, d = 200, 80, 3
    K = rng.uniform(-10, 10, (nk, d))
    Sk = rng.uniform(0, 2, nk)
    Sk[::5] = 0
    V = rng.uniform(0, 1, nk)
    p = rng.uniform(0, 1, N * d)
    pot = P.NUwNU(np.ones(N), K.ravel(), Sk, V, np.ones(d), method="direct")
    S0 = torch.tensor(Sk, requires_grad=True)
    Vt = torch.tensor(V, requires_grad=True)
    x = torch.tensor(p, requires_grad=True)
    E = A.structure_factor_loss(pot, x, S0, Vt)
    E.backward()
    ref = O.NUwNU(np.ones(N), K.ravel(), Sk, V, np.ones(d))
    E0, g0 = ref.get_energy_gradient(p)
    dS0, dV = O.target_grads(ref.rho(p), Sk, V, N)
    assert abs(E.item() - E0) / E0 < 1e-11
    assert rel(x.grad.numpy(), g0 * TWO_PI) < 1e-11
    assert rel(S0.grad.numpy(), dS0) < 1e-11
    assert rel(Vt.grad.numpy(), dV) < 1e-11


def test_set_targets_switches_on_new_modes(mods):
    # a mode outside the constructor's active set (V = 0 at k and -k) is
    # switched on: the pair tables are rebuilt
    A, P = mods
    rng = np.random.default_rng(7)
    d, n, N = 2, 9, 100
    M = n * n
    V0 = np.zeros(M)
    V0[5] = 1.0
    Sk = rng.uniform(0.5, 1.5, M)
    p = rng.uniform(0, 1, N * d)
    pot = P.NUwU(np.ones(N), 0, Sk, V0, np.ones(d), 1e-12)
    V1 = rng.uniform(0, 1, M)
    pot.set_targets(Sk, V1)
    E, g = pot.get_energy_gradient(p)
    E0, g0 = O.NUwU(np.ones(N), 0, Sk, V1, np.ones(d)).get_energy_gradient(p)
    assert abs(E - E0) / E0 < 1e-9
    assert rel(g, g0) < 1e-9
    import pickle

    E2 = pickle.loads(pickle.dumps(pot)).get_energy(p)
    assert abs(E2 - E0) / E0 < 1e-9


def _small_case(prec, seed=31):
    rng = np.random.default_rng(seed)
    d, n, N = 2, 17, 300
    M = n ** d
    radii = rng.uniform(0.6, 1.4, N)
    Sk = rng.uniform(0.2, 2, M)
    Sk[::3] = 0
    V = rng.uniform(0, 1, M)
    V[::7] = 0
    L = np.array([1.0, 1.5])
    return radii, Sk, V, L, rng, N, d


@pytest.mark.parametrize("prec,eps", [("fp64-adjoint", 1e-12), ("fp32", 1e-6), ("fp64", 1e-12)])
def test_split_evaluation_equals_fused_pass(mods, prec, eps):
    """energy_keep (spreading, transform, loss, pullback) followed by
    gradient_kept (inverse transform, interpolation) gives bitwise the
    energy and gradient of the fused pass; any other evaluation or a target
    update drops the kept state."""
    A, P = mods
    radii, Sk, V, L, rng, N, d = _small_case(prec)
    pot = P.NUwU(radii, 0, Sk, V, L, eps, precision=prec)
    x = torch.tensor(rng.uniform(0, 1, N * d), device="cuda")
    E0, g0 = pot.get_energy_gradient(x)
    tok, E1 = pot.energy_keep(x)
    assert tok and E1 == E0
    g1 = torch.empty_like(x)
    assert pot.gradient_kept(tok, g1) and torch.equal(g1, g0)
    assert pot.gradient_kept(tok)
    pot.get_energy(x)
    assert not pot.gradient_kept(tok)
    tok2, _ = pot.energy_keep(x)
    assert tok2 != tok and not pot.gradient_kept(tok) and pot.gradient_kept(tok2)
    pot.set_targets(Sk, V)
    assert not pot.gradient_kept(tok2)


def test_autograd_lazy_backward_and_interleaved_evaluations(mods):
    """Forward keeps the pass half done; backward finishes it.  With another
    evaluation of the same potential, or a target update, between the two,
    backward evaluates again at the forward's coordinates and targets (and
    leaves the potential's current targets in place).  Gradients against
    the oracle each time."""
    A, P = mods
    radii, Sk, V, L, rng, N, d = _small_case("fp64-adjoint", seed=41)
    pot = P.NUwU(radii, 0, Sk, V, L, 1e-12, precision="fp64-adjoint")
    scale = np.tile(TWO_PI / L, N)

    def oracle(p, s, v):
        ref = O.NUwU(radii, 0, s, v, L)
        E0, g0 = ref.get_energy_gradient(p)
        dS0, dV = O.target_grads(ref.rho(p), s, v, N)
        return E0, g0 * scale, dS0, dV

    p1, p2 = rng.uniform(0, 1, N * d), rng.uniform(0, 1, N * d)
    Sk2 = Sk * rng.uniform(0.5, 1.5, Sk.size)
    # 1: plain forward / backward
    x1 = torch.tensor(p1, device="cuda", requires_grad=True)
    S0 = torch.tensor(Sk, requires_grad=True)
    Vt = torch.tensor(V, requires_grad=True)
    E = A.structure_factor_loss(pot, x1, S0, Vt)
    E.backward()
    E0, gx0, dS0, dV = oracle(p1, Sk, V)
    assert abs(E.item() - E0) / abs(E0) < 1e-9
    assert rel(x1.grad.cpu(), gx0) < 1e-9 and rel(S0.grad, dS0) < 1e-9 and rel(Vt.grad, dV) < 1e-9
    # 2: two forwards, backward of the first after the second (re-evaluated)
    xa = torch.tensor(p1, device="cuda", requires_grad=True)
    xb = torch.tensor(p2, device="cuda", requires_grad=True)
    Ea = A.structure_factor_loss(pot, xa, S0, Vt)
    Eb = A.structure_factor_loss(pot, xb, S0, Vt)
    Ea.backward()
    Eb.backward()
    assert rel(xa.grad.cpu(), gx0) < 1e-9
    assert rel(xb.grad.cpu(), oracle(p2, Sk, V)[1]) < 1e-9
    # 3: a target update between forward and backward
    xc = torch.tensor(p1, device="cuda", requires_grad=True)
    S0c = torch.tensor(Sk, requires_grad=True)
    Ec = A.structure_factor_loss(pot, xc, S0c, Vt)
    pot.set_targets(Sk2, V)
    Ec.backward()
    assert rel(xc.grad.cpu(), gx0) < 1e-9 and rel(S0c.grad, dS0) < 1e-9
    E2 = oracle(p1, Sk2, V)[0]
    assert abs(pot.get_energy(p1) - E2) / abs(E2) < 1e-9  # the update is still in place
    # 4: the same target tensors again, then S0 changed in place (V kept)
    S0d = torch.tensor(Sk, requires_grad=True)
    for it in range(3):
        xd = torch.tensor(p1, device="cuda", requires_grad=True)
        Ed = A.structure_factor_loss(pot, xd, S0d, Vt)
        Ed.backward()
        s_now = S0d.detach().numpy().copy()
        Eo, go, dSo, _ = oracle(p1, s_now, V)
        assert abs(Ed.item() - Eo) / abs(Eo) < 1e-9
        assert rel(xd.grad.cpu(), go) < 1e-9 and rel(S0d.grad, dSo) < 1e-9
        S0d.grad = None
        with torch.no_grad():
            S0d.mul_(1.1)
    # 5: forward without backward, then a plain evaluation: nothing left over
    Ef = A.structure_factor_loss(pot, torch.tensor(p2, device="cuda", requires_grad=True), S0, Vt)
    assert abs(Ef.item() - oracle(p2, Sk, V)[0]) / abs(Ef.item()) < 1e-9
    assert abs(pot.get_energy(p1) - E0) / abs(E0) < 1e-9
