"""GPU parity of NUwU (libfresco_hip through the C ABI) against the oracle.

Tolerances (relative l2 unless stated):
  fp64 / fp64-adjoint at eps=1e-12 : energy 1e-9, gradient 1e-9, rho 1e-9
  fp32 (adjoint)      at eps=1e-6  : energy 1e-5, gradient 1e-5, rho 1e-5
"""
import math

import numpy as np
import pytest

from oracle import sf_direct as O

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-9, "fp64-adjoint": 1e-9, "fp32": 1e-5}
EPS = {"fp64": 1e-12, "fp64-adjoint": 1e-12, "fp32": 1e-6}


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def make_case(d, n, N, seed, spread_coords=True):
    rng = np.random.default_rng(seed)
    M = n ** d
    radii = rng.uniform(0.5, 1.5, N)
    Sk = rng.uniform(0, 2, M)
    Sk[::3] = 0.0
    V = rng.uniform(0, 1, M)
    V[::5] = 0.0
    lo, hi = (-0.3, 1.3) if spread_coords else (0.0, 1.0)
    p = rng.uniform(lo, hi, N * d)
    return radii, Sk, V, p


@pytest.fixture(scope="module")
def NUwU():
    from fresco_amd.potentials import NUwU as cls

    return cls


# (3, 21, 3000): a fine grid (42-48 cells) with few register-spreader tiles,
# so each tile's points are sliced over several wavefronts
@pytest.mark.parametrize("prec", ["fp64", "fp64-adjoint", "fp32"])
@pytest.mark.parametrize("d,n,N", [(1, 33, 200), (2, 17, 400), (3, 9, 300), (3, 21, 3000)])
def test_parity_vs_direct_sum(NUwU, prec, d, n, N):
    radii, Sk, V, p = make_case(d, n, N, seed=10 * d + n)
    L = np.ones(d)
    ref = O.NUwU(radii, 0.0, Sk, V, L)
    E0, g0 = ref.get_energy_gradient(p)
    pot = NUwU(radii, 0.0, Sk, V, L, EPS[prec], precision=prec)
    E, g = pot.get_energy_gradient(p)
    print("case", d, n, N, prec, pot.plan_info(), flush=True)
    assert abs(E - E0) / abs(E0) < TOL[prec]
    assert rel(g, g0) < TOL[prec]
    assert rel(pot.get_rho(), ref.rho(p)) < TOL[prec]
    Ee = pot.get_energy(p)
    assert abs(Ee - ref.get_energy(p)) / abs(E0) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_box_not_one(NUwU, prec):
    # the reference map (p - round(p/L)) * 2pi/L is kept for L != 1 (nuwu.hpp:219)
    d = 2
    radii, Sk, V, p = make_case(d, 13, 150, seed=5)
    L = np.array([2.0, 0.5])
    ref = O.NUwU(radii, 0.0, Sk, V, L)
    E0, g0 = ref.get_energy_gradient(p)
    pot = NUwU(radii, 0.0, Sk, V, L, EPS[prec], precision=prec)
    E, g = pot.get_energy_gradient(p)
    assert abs(E - E0) / abs(E0) < TOL[prec]
    assert rel(g, g0) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp64-adjoint", "fp32"])
def test_square_lattice_closed_form(NUwU, prec):
    # points exactly on a lattice (and on fine-grid cells): E = count * N^2
    n = 8
    g = (np.arange(n) + 0.5) / n
    p = np.stack(np.meshgrid(g, g, indexing="ij"), axis=-1).ravel()
    N = n * n
    kp = 12
    M = (2 * kp + 1) ** 2
    pot = NUwU(np.ones(N), 0, np.zeros(M), np.ones(M), np.ones(2), EPS[prec], precision=prec)
    count = sum(1 for a in range(-kp, kp + 1) for b in range(-kp, kp + 1) if a % n == 0 and b % n == 0)
    E, grad = pot.get_energy_gradient(p)
    assert abs(E - count * N ** 2) / (count * N ** 2) < TOL[prec]
    # a perfect lattice is a stationary point: the gradient is zero up to the
    # path's error relative to the gradient scale of a displaced lattice
    rng = np.random.default_rng(1)
    q = p + rng.normal(scale=0.02 / n, size=p.size)
    _, g0 = O.NUwU(np.ones(N), 0, np.zeros(M), np.ones(M), np.ones(2)).get_energy_gradient(q)
    assert np.linalg.norm(grad) < TOL[prec] * np.linalg.norm(g0)


def test_deterministic(NUwU):
    radii, Sk, V, p = make_case(3, 9, 500, seed=11)
    for prec in ("fp32", "fp64"):
        pot = NUwU(radii, 0.0, Sk, V, np.ones(3), EPS[prec], precision=prec)
        E1, g1 = pot.get_energy_gradient(p)
        E2, g2 = pot.get_energy_gradient(p)
        assert E1 == E2
        assert np.array_equal(g1, g2)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_stream_ordered_evaluation_matches_synchronous(NUwU, prec):
    """energy_gradient_device returns once the energy is on the host (the
    gradient completes on the stream, read after it): back-to-back
    evaluations queue behind each other, and their energies and gradients
    equal the host-synchronous C-ABI call's bitwise (3D, band form for FP32)."""
    import ctypes

    import torch

    from fresco_amd import _lib
    from fresco_amd._device import stream_handle

    radii, Sk, V, p = make_case(3, 21, 6000, seed=17)
    pot = NUwU(radii, 0.0, Sk, V, np.ones(3), EPS[prec], precision=prec)
    x = torch.tensor(p, device="cuda")
    gs = [torch.empty_like(x) for _ in range(6)]
    es = [pot.energy_gradient_device(x, g) for g in gs]
    g0 = torch.empty_like(x)
    e0 = ctypes.c_double()
    _lib.check(_lib.load().fresco_pot_energy_gradient(
        pot._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(g0.data_ptr()), ctypes.byref(e0), stream_handle()))
    torch.cuda.synchronize()
    assert all(e == e0.value for e in es)
    assert all(torch.equal(g, g0) for g in gs)
    E, g = O.NUwU(radii, 0.0, Sk, V, np.ones(3)).get_energy_gradient(p)
    assert abs(e0.value - E) / abs(E) < TOL[prec]
    assert rel(g0.cpu().numpy(), g) < TOL[prec]


def test_torch_tensor_in_device_tensor_out(NUwU):
    import torch

    radii, Sk, V, p = make_case(2, 11, 100, seed=12)
    pot = NUwU(radii, 0.0, Sk, V, np.ones(2), 1e-10, precision="fp64")
    xt = torch.tensor(p, device="cuda")
    E, g = pot.get_energy_gradient(xt)
    assert isinstance(g, torch.Tensor) and g.is_cuda
    E2, g2 = pot.get_energy_gradient(p)
    assert E == E2
    assert np.array_equal(g.cpu().numpy(), g2)


def test_invalid_sizes_and_faults(NUwU):
    from fresco_amd import _lib

    with pytest.raises(ValueError):
        NUwU(np.ones(10), 0, np.zeros(64), np.ones(64), np.ones(2))  # 8^2 is not (2K+1)^2
    with pytest.raises(ValueError):
        NUwU(np.ones(10), 0, np.zeros(9), np.ones(9), np.ones(2), precision="fp16")
    pot = NUwU(np.ones(10), 0, np.zeros(9), np.ones(9), np.ones(2))
    with pytest.raises(ValueError):
        pot.get_energy(np.zeros(7))
    _lib.check(_lib.load().fresco_set_fault(1))
    with pytest.raises(_lib.FrescoError):
        NUwU(np.ones(10), 0, np.zeros(9), np.ones(9), np.ones(2))
    _lib.check(_lib.load().fresco_set_fault(0))
    NUwU(np.ones(10), 0, np.zeros(9), np.ones(9), np.ones(2))


def test_fd_directional_derivative_fp64(NUwU):
    # mid-size 3D case, fp64: (E(p + h v) - E(p - h v)) / 2h == grad . v * 2pi
    rng = np.random.default_rng(13)
    N, n, d = 4000, 33, 3
    M = n ** d
    V = (np.linalg.norm(O.mode_grid(16, 3), axis=1) <= 12).astype(float)
    V[M // 2] = 0
    pot = NUwU(np.ones(N), 0, np.zeros(M), V, np.ones(d), 1e-12, precision="fp64")
    p = rng.uniform(0, 1, N * d)
    v = rng.normal(size=N * d)
    h = 1e-6
    E, g = pot.get_energy_gradient(p)
    fd = (pot.get_energy(p + h * v) - pot.get_energy(p - h * v)) / (2 * h)
    assert abs(fd - 2 * math.pi * np.dot(g, v)) / abs(fd) < 1e-6


def test_3d_n2p20_fp32_vs_fp64_and_invariants(NUwU):
    """Stealthy |k| <= 24 on 129^3 modes, N = 2^20: FP32-adjoint vs FP64 plain
    (energy and gradient within 1e-5), Sum(grad) = 0 and translation invariance
    of the energy."""
    import torch

    rng = np.random.default_rng(0)
    N, n, d = 1 << 20, 129, 3
    k = O.mode_grid(64, 3)
    kk = np.linalg.norm(k, axis=1)
    V = ((kk > 0) & (kk <= 24)).astype(float)
    Sk = np.zeros_like(V)
    p = rng.uniform(0, 1, N * d)
    p64 = NUwU(np.ones(N), 64, Sk, V, np.ones(d), 1e-9, precision="fp64")
    p32 = NUwU(np.ones(N), 64, Sk, V, np.ones(d), 1e-6, precision="fp32")
    xt = torch.tensor(p, device="cuda")
    E64, g64 = p64.get_energy_gradient(xt)
    E32, g32 = p32.get_energy_gradient(xt)
    print("3D N=2^20 energies", E64, E32, flush=True)
    assert abs(E32 - E64) / abs(E64) < 1e-5
    r = float(torch.linalg.norm(g32 - g64) / torch.linalg.norm(g64))
    print("3D N=2^20 gradient rel err fp32 vs fp64", r, flush=True)
    assert r < 1e-5
    gs = g64.reshape(-1, 3).sum(0).abs().max().item()
    assert gs < 1e-8 * float(g64.abs().max()) * N ** 0.5
    Es = p64.get_energy(xt + 0.123456)
    assert abs(Es - E64) / abs(E64) < 1e-7


@pytest.mark.parametrize("prec", ["fp64-adjoint", "fp32"])
def test_active_band_grid(NUwU, prec):
    """V = 0 outside |k| <= 4 on a 21^3 mode grid: the adjoint paths size the
    fine grid by the active band; energy / gradient / full-grid rho and the
    target derivatives still match the exact sums, and switching on modes
    beyond the band rebuilds the grid."""
    rng = np.random.default_rng(31)
    d, n, N = 3, 21, 2000
    k = O.mode_grid(10, 3)
    kk = np.linalg.norm(k, axis=1)
    V = ((kk > 0) & (kk <= 4)).astype(float)
    Sk = rng.uniform(0.5, 2.0, n ** d)
    p = rng.uniform(0, 1, N * d)
    radii = rng.uniform(0.8, 1.2, N)
    pot = NUwU(radii, 0, Sk, V, np.ones(d), EPS[prec], precision=prec)
    full = NUwU(radii, 0, Sk, V, np.ones(d), EPS[prec], precision="fp64")
    assert pot.plan_info()["nf"][0] < full.plan_info()["nf"][0]
    ref = O.NUwU(radii, 0, Sk, V, np.ones(d))
    E0, g0 = ref.get_energy_gradient(p)
    E, g = pot.get_energy_gradient(p)
    assert abs(E - E0) / abs(E0) < TOL[prec]
    assert rel(g, g0) < TOL[prec]
    assert rel(pot.get_rho(), ref.rho(p)) < TOL[prec]
    dS0, dV = pot.target_grads()
    eS0, eV = O.target_grads(ref.rho(p), Sk, V, N)
    assert rel(dS0.cpu().numpy(), eS0) < TOL[prec]
    assert rel(dV.cpu().numpy(), eV) < TOL[prec]
    V2 = ((kk > 0) & (kk <= 7)).astype(float)
    pot.set_targets(Sk, V2)
    E2, g2 = pot.get_energy_gradient(p)
    r2 = O.NUwU(radii, 0, Sk, V2, np.ones(d))
    E20, g20 = r2.get_energy_gradient(p)
    assert abs(E2 - E20) / abs(E20) < TOL[prec]
    assert rel(g2, g20) < TOL[prec]


@pytest.mark.parametrize("prec,eps,tol", [("fp32", 1e-6, 1e-5), ("fp64", 1e-7, 1e-6)])
def test_halo_spreader_parity_and_determinism(NUwU, prec, eps, tol):
    """3D grids of few bins with a kernel at most 8 wide take the input-driven
    halo spreader (LDS tile plus one-sided halo, fixed-order merge): parity
    with the exact sums and bitwise-equal repeated evaluations."""
    radii, Sk, V, p = make_case(3, 21, 5000, seed=41)
    pot = NUwU(radii, 0.0, Sk, V, np.ones(3), eps, precision=prec)
    info = pot.plan_info()
    assert info["w"] <= 8 and min(info["nf"]) >= 23
    ref = O.NUwU(radii, 0.0, Sk, V, np.ones(3))
    E0, g0 = ref.get_energy_gradient(p)
    E, g = pot.get_energy_gradient(p)
    assert abs(E - E0) / abs(E0) < tol
    assert rel(g, g0) < tol
    E2, g2 = pot.get_energy_gradient(p)
    assert E == E2 and np.array_equal(g, g2)


@pytest.mark.parametrize("prec,eps,tol", [("fp32", 1e-6, 1e-5), ("fp64", 1e-7, 1e-6)])
def test_halo_spreader_wavefront_layouts_agree(NUwU, prec, eps, tol, monkeypatch):
    """The halo spreader with one wavefront per tile and with four (z-planes
    split by residue mod 4), at one and three point slices per bin: parity
    with the exact sums, and the two layouts bitwise equal at equal slices
    (same products, same per-cell order)."""
    radii, Sk, V, p = make_case(3, 21, 5000, seed=43)
    ref = O.NUwU(radii, 0.0, Sk, V, np.ones(3))
    E0, g0 = ref.get_energy_gradient(p)
    res = {}
    for mode in ("halo:1", "halo:3", "halo4:1", "halo4:3"):
        monkeypatch.setenv("FRESCO_SPREAD3", mode)
        pot = NUwU(radii, 0.0, Sk, V, np.ones(3), eps, precision=prec)
        info = pot.plan_info()
        assert info["spreader"] == mode.split(":")[0] and info["spread_parts"] == int(mode[-1])
        E, g = pot.get_energy_gradient(p)
        assert abs(E - E0) / abs(E0) < tol
        assert rel(g, g0) < tol
        res[mode] = (E, g)
    for sl in ("1", "3"):
        assert res["halo:" + sl][0] == res["halo4:" + sl][0]
        assert np.array_equal(res["halo:" + sl][1], res["halo4:" + sl][1])


def _clustered(N, frac, seed):
    """N points, a fraction frac of them in a cube of side 0.04 (a few bins
    then hold hundreds to thousands of points, most bins none)."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(0, 1, (N, 3))
    nc = int(frac * N)
    p[:nc] = 0.3 + 0.04 * rng.uniform(0, 1, (nc, 3))
    return p.ravel()


@pytest.mark.parametrize("case", ["uniform", "clustered"])
def test_matrix_core_spreader(NUwU, case, monkeypatch):
    """The FP32 halo spreading as matrix products, at one and two slices per
    bin: points ranked by the z of their first cell in each x window, the
    window's planes accumulated in registers (halom), and points ranked by
    the y and z of their first cell, the y rows kept in a ring and written
    row by row, one tile per window (halor).  Parity with the exact sums,
    agreement with the LDS halo spreader to FP32 rounding, bitwise repeats.
    The clustered case puts over a thousand points into single bins (several
    ranking rounds per slice, the later rounds adding to the tiles, runs
    padded to whole groups) and leaves most bins empty."""
    n, N = 21, 6000
    radii, Sk, V, p = make_case(3, n, N, seed=81)
    if case == "clustered":
        p = _clustered(N, 0.7, seed=82)
    ref = O.NUwU(radii, 0.0, Sk, V, np.ones(3))
    E0, g0 = ref.get_energy_gradient(p)
    res = {}
    for mode in ("halom:2", "halom:4", "halor:2", "halor:4", "halo4:2"):
        monkeypatch.setenv("FRESCO_SPREAD3", mode)
        pot = NUwU(radii, 0.0, Sk, V, np.ones(3), 1e-6, precision="fp32")
        info = pot.plan_info()
        print(case, mode, info, flush=True)
        assert info["spreader"] == mode.split(":")[0] and info["spread_parts"] == int(mode[-1])
        E, g = pot.get_energy_gradient(p)
        assert abs(E - E0) / abs(E0) < 1e-5
        assert rel(g, g0) < 1e-5
        assert rel(pot.get_rho(), ref.rho(p)) < 1e-5
        E2, g2 = pot.get_energy_gradient(p)
        assert E == E2 and np.array_equal(g, g2)
        res[mode] = (E, g)
    for mode in ("halom:2", "halom:4", "halor:2", "halor:4"):
        assert abs(res[mode][0] - res["halo4:2"][0]) / abs(E0) < 2e-6
        assert rel(res[mode][1], res["halo4:2"][1]) < 2e-6


def test_ring_spreader_with_many_bins(NUwU, monkeypatch):
    """A plan with more than 16384 bins (every mode constrained: the full
    129^3 mode grid), where the halo buffer holds a single halo tile per bin:
    the ring spreader's two window tiles per bin fit it, and the gradient
    agrees with the four-wavefront LDS spreader's to FP32 rounding."""
    n, N = 129, 20000
    rng = np.random.default_rng(83)
    p = rng.uniform(0, 1, 3 * N)
    M = n ** 3
    Sk = np.zeros(M)
    V = np.ones(M)
    res = {}
    for mode in ("halor:2", "halo4:1"):
        monkeypatch.setenv("FRESCO_SPREAD3", mode)
        pot = NUwU(np.ones(N), 0.0, Sk, V, np.ones(3), 1e-6, precision="fp32")
        info = pot.plan_info()
        print(mode, info, flush=True)
        nb = [-(-info["nf"][a] // info["tile"][a]) for a in range(3)]
        assert nb[0] * nb[1] * nb[2] > 16384
        assert (info["spreader"], info["spread_parts"]) == (mode.split(":")[0], int(mode[-1]))
        res[mode] = pot.get_energy_gradient(p)
    E0, g0 = res["halo4:1"]
    assert abs(res["halor:2"][0] - E0) / abs(E0) < 2e-6
    assert rel(res["halor:2"][1], g0) < 2e-6


@pytest.mark.parametrize("case", ["uniform", "clustered"])
def test_interpolation_lane_layouts_agree(NUwU, case, monkeypatch):
    """3D FP32 gradient interpolation from 16-byte staged regions with four
    lanes per point (two x offsets each, quad broadcasts and reductions) and
    with eight (FRESCO_INTERP_LPP=8), at one and two slices per bin: parity
    with the exact sums, agreement to FP32 rounding, bitwise repeats.  The
    clustered case gives slices of more than 1024 points (interpolated in
    bin order, not sorted)."""
    n, N = 33, 6000
    radii, Sk, V, p = make_case(3, n, N, seed=83)
    if case == "clustered":
        p = _clustered(N, 0.7, seed=84)
    ref = O.NUwU(radii, 0.0, Sk, V, np.ones(3))
    E0, g0 = ref.get_energy_gradient(p)
    res = {}
    for lpp, wgs in (("4", "2048"), ("4", "1"), ("8", "2048")):
        monkeypatch.setenv("FRESCO_INTERP_LPP", lpp)
        monkeypatch.setenv("FRESCO_INTERP_WGS", wgs)
        pot = NUwU(radii, 0.0, Sk, V, np.ones(3), 1e-6, precision="fp32")
        info = pot.plan_info()
        assert info["nf"][0] % 4 == 0 and info["tile"] == [16, 8, 8], info
        E, g = pot.get_energy_gradient(p)
        assert abs(E - E0) / abs(E0) < 1e-5
        assert rel(g, g0) < 1e-5
        E2, g2 = pot.get_energy_gradient(p)
        assert E == E2 and np.array_equal(g, g2)
        res[lpp + ":" + wgs] = g
    assert rel(res["4:2048"], res["8:2048"]) < 2e-6
    assert rel(res["4:1"], res["4:2048"]) < 2e-6


@pytest.mark.parametrize("case", ["uniform", "clustered"])
def test_matrix_core_spreader_2d(NUwU, case, monkeypatch):
    """2D FP32 halo spreading as matrix products (16 x 16 bins: four 16 x 16
    windows of the 23 x 23 halo tile, one product per 4 points of a window,
    windows summed into the tile in a fixed order): parity with the exact
    sums, agreement with the LDS spreader to FP32 rounding, bitwise repeats.
    Clustered points put hundreds to thousands of points into single bins
    (several ranking windows per slice)."""
    n, N = 65, 8000
    radii, Sk, V, p = make_case(2, n, N, seed=91)
    if case == "clustered":
        rng = np.random.default_rng(92)
        q = rng.uniform(0, 1, (N, 2))
        q[: int(0.7 * N)] = 0.4 + 0.03 * rng.uniform(0, 1, (int(0.7 * N), 2))
        p = q.ravel()
    ref = O.NUwU(radii, 0.0, Sk, V, np.ones(2))
    E0, g0 = ref.get_energy_gradient(p)
    res = {}
    for mode in ("mfma", "halo"):
        if mode == "halo":
            monkeypatch.setenv("FRESCO_SPREAD2", "halo")
        else:
            monkeypatch.delenv("FRESCO_SPREAD2", raising=False)
        pot = NUwU(radii, 0.0, Sk, V, np.ones(2), 1e-6, precision="fp32")
        info = pot.plan_info()
        print(case, mode, info, flush=True)
        assert info["tile"][:2] == [16, 16]
        E, g = pot.get_energy_gradient(p)
        assert abs(E - E0) / abs(E0) < 1e-5
        assert rel(g, g0) < 1e-5
        E2, g2 = pot.get_energy_gradient(p)
        assert E == E2 and np.array_equal(g, g2)
        res[mode] = (E, g)
    assert abs(res["mfma"][0] - res["halo"][0]) / abs(E0) < 2e-6
    assert rel(res["mfma"][1], res["halo"][1]) < 2e-6


@pytest.mark.parametrize(
    "d,n,N", [(3, 33, 5000), (3, 21, 2047), (3, 41, 300007), (2, 65, 4097), (1, 129, 3000)]
)
def test_direct_binning_equals_radix_sort(NUwU, d, n, N, monkeypatch):
    """Plans whose bins fit one 11-bit digit are binned by one counting pass
    (tile counts, group prefixes, ballot-ranked placement) instead of the
    radix sort; both give index order inside every bin, so energy and
    gradient are bitwise equal.  Sizes off the 1024-point tile and its
    256-point wavefront quarters, up to 293 tiles in several groups."""
    radii, Sk, V, p = make_case(d, n, N, seed=60 + d)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("FRESCO_DIRECT_BINNING", mode)
        monkeypatch.setenv("FRESCO_SPREAD3", "halo4:2")
        pot = NUwU(radii, 0.0, Sk, V, np.ones(d), 1e-6, precision="fp32")
        out[mode] = pot.get_energy_gradient(p)
    assert out["1"][0] == out["0"][0]
    assert np.array_equal(out["1"][1], out["0"][1])


@pytest.mark.parametrize("kmax", [6, 12])
def test_band_dft_matches_fft_path(NUwU, kmax, monkeypatch):
    """3D FP32 adjoint path with an active band |k_a| <= B small against the
    fine grid: its spectrum comes from pruned separable DFTs of the band
    instead of the grid's real FFTs, in the fused form (x and y passes per
    plane, the z pass with the loss and the inverse z pass per tile of band
    columns, y and x passes per plane) and in the four-launch form
    (FRESCO_BANDDFT_FUSED=0).  Same energy (with and without the gradient),
    gradient and target derivatives as the FFT path (FRESCO_BANDDFT=0) to
    FP32 rounding, and parity with the exact sums."""
    rng = np.random.default_rng(50 + kmax)
    n, N = 33, 4000
    k = O.mode_grid(n // 2, 3)
    kk = np.linalg.norm(k, axis=1)
    V = np.where((kk > 0) & (kk <= kmax), rng.uniform(0.2, 1.0, len(kk)), 0.0)
    Sk = np.where(rng.uniform(size=len(kk)) < 0.5, 0.0, rng.uniform(0.5, 2.0, len(kk)))
    radii = rng.uniform(0.7, 1.3, N)
    p = rng.uniform(0, 1, 3 * N)
    out = {}
    for mode, fused in (("1", "1"), ("1", "0"), ("0", "1")):
        monkeypatch.setenv("FRESCO_BANDDFT", mode)
        monkeypatch.setenv("FRESCO_BANDDFT_FUSED", fused)
        pot = NUwU(radii, 0.0, Sk, V, np.ones(3), 1e-6, precision="fp32")
        E, g = pot.get_energy_gradient(p)
        Ee = pot.get_energy(p)
        dS0, dV = pot.target_grads()
        out[mode + fused] = (E, g, dS0.cpu().numpy(), dV.cpu().numpy(), Ee)
    (E1, g1, s1, v1, e1) = out["11"]
    for key in ("10", "01"):
        E0, g0, s0, v0, e0 = out[key]
        assert abs(E1 - E0) / abs(E0) < 2e-6, key
        assert abs(e1 - e0) / abs(E0) < 2e-6, key
        assert rel(g1, g0) < 2e-6, key
        assert rel(s1, s0) < 2e-6 and rel(v1, v0) < 2e-6, key
    ref = O.NUwU(radii, 0.0, Sk, V, np.ones(3))
    Er, gr = ref.get_energy_gradient(p)
    assert abs(E1 - Er) / abs(Er) < 1e-5
    assert rel(g1, gr) < 1e-5


@pytest.mark.parametrize("band", [0, 1, None])
def test_band_dft_smallest_bands(NUwU, band, monkeypatch):
    """The band transform at its smallest: weight only on k = 0 (band 0), on
    |k_a| <= 1 (band 1), and on no mode at all (energy and gradient 0);
    parity with the exact sums and with the FFT path."""
    rng = np.random.default_rng(70)
    n, N = 9, 500
    k = O.mode_grid(n // 2, 3)
    V = np.zeros(len(k))
    if band is not None:
        V[np.max(np.abs(k), axis=1) <= band] = rng.uniform(0.5, 1.0, int(np.sum(np.max(np.abs(k), axis=1) <= band)))
    Sk = rng.uniform(0.0, 2.0, len(k))
    radii = rng.uniform(0.8, 1.2, N)
    p = rng.uniform(0, 1, 3 * N)
    res = {}
    for mode, fused in (("1", "1"), ("1", "0"), ("0", "1")):
        monkeypatch.setenv("FRESCO_BANDDFT", mode)
        monkeypatch.setenv("FRESCO_BANDDFT_FUSED", fused)
        pot = NUwU(radii, 0.0, Sk, V, np.ones(3), 1e-6, precision="fp32")
        res[mode + fused] = pot.get_energy_gradient(p)
    for key in ("11", "10"):
        E, g = res[key]
        if band is None:
            assert E == 0.0 and not np.any(g)
            continue
        Er, gr = O.NUwU(radii, 0.0, Sk, V, np.ones(3)).get_energy_gradient(p)
        assert abs(E - Er) / abs(Er) < 1e-5
        assert abs(E - res["01"][0]) / abs(Er) < 2e-6
        gn = max(np.linalg.norm(gr), 1e-300)
        assert np.linalg.norm(g - gr) / gn < 1e-5 or np.linalg.norm(g - gr) < 1e-5 * abs(Er)


@pytest.mark.parametrize("prec", ["fp32", "fp64-adjoint"])
def test_adjoint_gradient_with_weighted_k0(NUwU, prec):
    """V(k = 0) != 0 on a small 1D grid: the k = 0 term carries rho_0 = N but
    moves no point; the adjoint path keeps it off the grid whose derivative it
    interpolates (with it, a constant of size N turned the kernel's ripple into
    gradient errors of 1e-5 in FP64 and 1e-4 in FP32)."""
    rng = np.random.default_rng(1)
    N, n = 60, 9
    radii = rng.uniform(0.6, 1.4, N)
    Sk = rng.uniform(0.3, 2, n)
    Sk[::3] = 0
    V = rng.uniform(0, 1, n)
    V[n // 2] = 1.0
    p = rng.uniform(-0.2, 1.2, N)
    E0, g0 = O.NUwU(radii, 0, Sk, V, np.ones(1)).get_energy_gradient(p)
    E, g = NUwU(radii, 0, Sk, V, np.ones(1), 1e-6, precision=prec).get_energy_gradient(p)
    assert abs(E - E0) / abs(E0) < 1e-5
    assert rel(g, g0) < (1e-5 if prec == "fp32" else 1e-6)
