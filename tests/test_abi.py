"""The C-ABI library loads without a GPU and exports every symbol that
include/fresco_hip.h declares (no compute calls)."""
import ctypes
import re

import pytest

from fresco_amd import _lib


def declared_functions():
    text = open(_lib.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fresco_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = declared_functions()
    assert "fresco_pot_energy_gradient" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_python_signatures_cover_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_version_and_error_without_gpu():
    lib = _lib.load()
    assert b"gfx950" in lib.fresco_version()
    # a null-handle call fails cleanly with an invalid-argument status
    e = ctypes.c_double()
    rc = lib.fresco_pot_energy(None, None, ctypes.byref(e), None)
    assert rc == _lib.FRESCO_E_INVALID
    with pytest.raises(ValueError):
        _lib.check(rc)


def test_reference_import_path():
    """The reference's import names (fresco/potentials/__init__.py:2-14,
    fresco/optimize/__init__.py:1-3, the examples' `from fresco.utils import
    export_structure`) resolve to the GPU implementations."""
    import fresco.distances
    import fresco.optimize
    import fresco.potentials
    import fresco.utils
    import fresco_amd.optimize
    import fresco_amd.potentials
    import fresco_amd.utils

    for name in ("UwU", "NUwU", "UwNU", "NUwNU", "InversePower", "WCAPotential", "BasePotential",
                 "CombinedPotential", "as_cpp_potential"):
        assert getattr(fresco.potentials, name) is getattr(fresco_amd.potentials, name), name
    for name in ("Result", "LBFGS_CPP", "ModifiedFireCPP"):
        assert getattr(fresco.optimize, name) is getattr(fresco_amd.optimize, name), name
    assert fresco.utils.export_structure is fresco_amd.utils.export_structure
    assert fresco.distances.Distance.PERIODIC.value == 2


def test_reference_optimizer_adapter_library_loads():
    """oracle/_ref/libref_hip_opt.so (the reference's lbfgs.cpp with
    include/fresco_hip_potential.hpp) links against this library: loading it
    maps the same libfresco_hip.so, so handles made here are valid there."""
    import os

    from oracle import ref_opt as RO

    if not RO.hip_available():
        pytest.skip("oracle/_ref/libref_hip_opt.so not built (needs the reference sources)")
    _lib.load()
    L = RO.hip_lib()
    for name in ("ref_hip_lbfgs_create", "ref_hip_fire_create", "ref_hip_one_iteration", "ref_hip_state"):
        assert hasattr(L, name)
    maps = open("/proc/self/maps").read()
    paths = {line.split()[-1] for line in maps.splitlines() if line.endswith("libfresco_hip.so")}
    assert paths == {os.path.realpath(_lib.LIB_PATH)}
