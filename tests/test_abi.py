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
