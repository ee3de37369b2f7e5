"""fresco_radial_profile_noalloc: the allocation-free variant of
fresco_radial_profile (the shell sums of radial_profile,
fresco/utils/utils.py:138-152), called through the C ABI.

Its buffers live in the library's fixed per-GPU scratch array and it runs
the same kernels in the same order as fresco_radial_profile, so its sums and
counts must equal that entry point's exactly, also on a second, smaller call
that finds the first call's data still in the scratch array.
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from fresco_amd import _lib  # noqa: E402

pytestmark = pytest.mark.gpu


def _run(name, data, center):
    """Call entry point `name` on the device grid `data`; returns (sums, counts) on the host."""
    shape = np.asarray(data.shape, dtype=np.int64)
    center = np.ascontiguousarray(center, dtype=np.float64)
    nb = int(shape[0]) // 2 + 2
    sums = torch.full((nb,), float("nan"), dtype=torch.float64, device=data.device)
    counts = torch.full((nb,), -1, dtype=torch.int64, device=data.device)
    rc = getattr(_lib.load(), name)(
        ctypes.c_void_p(data.data_ptr()), data.dim(), shape.ctypes.data_as(_lib.c_int64_p), _lib.dptr(center),
        ctypes.c_void_p(sums.data_ptr()), ctypes.c_void_p(counts.data_ptr()), None,
    )
    return rc, sums.cpu().numpy(), counts.cpu().numpy()


def _grid(shape, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(shape, generator=g, dtype=torch.float64).to("cuda:0").contiguous()


def test_noalloc_equals_allocating_entry_point_on_consecutive_calls():
    torch.cuda.set_device(0)
    # first call: 3D, 864,000 cells (near the 2^20 limit), fractional centre
    # second call: a smaller 2D grid over the first call's leftovers
    # third call: 1D, smaller still
    cases = [((96, 100, 90), (47.5, 50.25, 44.0), 1), ((300, 257), (150.5, 128.0), 2), ((777,), (388.0,), 3)]
    for shape, center, seed in cases:
        print("noalloc radial profile: shape", shape, "centre", center, flush=True)
        data = _grid(shape, seed)
        rc_a, s_a, c_a = _run("fresco_radial_profile", data, center)
        rc_b, s_b, c_b = _run("fresco_radial_profile_noalloc", data, center)
        assert rc_a == _lib.FRESCO_OK, _lib.load().fresco_last_error()
        assert rc_b == _lib.FRESCO_OK, _lib.load().fresco_last_error()
        assert np.array_equal(c_a, c_b), shape
        assert c_b.sum() == data.numel()
        # same kernels, same summation order: equal sums, bit for bit
        assert np.array_equal(s_a, s_b), (shape, np.max(np.abs(s_a - s_b)))


def test_noalloc_matches_numpy_shell_sums():
    torch.cuda.set_device(0)
    shape, center = (64, 48), (31.5, 20.0)
    data = _grid(shape, 7)
    rc, s, c = _run("fresco_radial_profile_noalloc", data, center)
    assert rc == _lib.FRESCO_OK, _lib.load().fresco_last_error()
    idx = np.indices(shape).astype(np.float64)
    r = np.floor(np.sqrt(sum((idx[a] - center[a]) ** 2 for a in range(2)))).astype(np.int64)
    R = shape[0] // 2
    r = np.minimum(r, R + 1).ravel()
    host = data.cpu().numpy().ravel()
    assert np.array_equal(c, np.bincount(r, minlength=R + 2))
    np.testing.assert_allclose(s, np.bincount(r, weights=host, minlength=R + 2), rtol=1e-12, atol=1e-12)


def test_noalloc_rejects_grids_beyond_its_limits():
    torch.cuda.set_device(0)
    too_many_cells = torch.zeros((1025, 1024), dtype=torch.float64, device="cuda:0")
    rc, _, _ = _run("fresco_radial_profile_noalloc", too_many_cells, (512.0, 512.0))
    assert rc == _lib.FRESCO_E_INVALID
    assert b"limit" in _lib.load().fresco_last_error()
    too_many_shells = torch.zeros((5000,), dtype=torch.float64, device="cuda:0")
    rc, _, _ = _run("fresco_radial_profile_noalloc", too_many_shells, (2500.0,))
    assert rc == _lib.FRESCO_E_INVALID
    assert b"shells" in _lib.load().fresco_last_error()
    # the allocating entry point still takes both
    rc, _, c = _run("fresco_radial_profile", too_many_shells, (2500.0,))
    assert rc == _lib.FRESCO_OK and c.sum() == 5000
