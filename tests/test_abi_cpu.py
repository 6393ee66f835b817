"""The C-ABI library loads and exports every entry point include/dgsem_amd.h
declares (no compute calls: this container has no GPU)."""
import ctypes
import os
import re

import pytest

from dgsem_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "dgsem_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dgsem_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert set(names) == set(_abi.EXPORTED), names


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


def test_abi_version_and_defaults_without_gpu():
    lib = _abi.load()
    assert lib.dgsem_abi_version() == 2
    c = _abi.Config()
    lib.dgsem_config_default(ctypes.byref(c), 3, 3)
    assert (c.dim, c.polydeg, c.gamma, c.alpha_max, c.alpha_min) == (3, 3, 1.4, 0.5, 0.001)
    assert c.mu == 6.25e-4 and c.prandtl == 0.72
    lib.dgsem_config_default(ctypes.byref(c), 2, 4)
    assert c.mu == 5.0e-3 and c.prandtl == 0.73


def test_invalid_configuration_is_rejected_with_a_message():
    lib = _abi.load()
    c = _abi.Config()
    lib.dgsem_config_default(ctypes.byref(c), 4, 3)  # dim 4 is invalid
    h = ctypes.c_void_p()
    rc = lib.dgsem_create(ctypes.byref(c), None, ctypes.byref(h))
    assert rc == -1
    assert b"dim" in lib.dgsem_last_error(h)
    lib.dgsem_destroy(h)


@pytest.mark.parametrize("dim,faces,rc,word", [
    (2, [1, 0, 0, 0], -1, b"periodic on both faces"),   # -x wall, +x periodic
    (2, [7, 7, 0, 0], -1, b"unknown boundary"),
    (3, [1, 1, 0, 0, 0, 0], -2, b"2D meshes"),
    (2, [4, 4, 0, 0], -1, b"Navier-Stokes boundary"),  # an adiabatic wall under Euler
])
def test_boundary_configuration_errors_without_gpu(dim, faces, rc, word):
    """Face kinds are checked before any device work: a direction must be
    periodic on both faces or on neither, kinds must be known and
    non-periodic faces are built for 2D Euler only."""
    lib = _abi.load()
    c = _abi.Config()
    lib.dgsem_config_default(ctypes.byref(c), dim, 3)
    for f, k in enumerate(faces):
        c.face_boundary[f] = k
    h = ctypes.c_void_p()
    assert lib.dgsem_create(ctypes.byref(c), None, ctypes.byref(h)) == rc
    assert word in lib.dgsem_last_error(h)
    lib.dgsem_destroy(h)


def test_direction_kind_applies_to_both_faces_without_gpu():
    """boundary[dim] sets both faces of a direction: a 3D mesh with x walls is
    refused as not built, which shows both x faces took the wall kind."""
    lib = _abi.load()
    c = _abi.Config()
    lib.dgsem_config_default(ctypes.byref(c), 3, 3)
    c.boundary[0] = _abi.BC_SLIP_WALL
    h = ctypes.c_void_p()
    assert lib.dgsem_create(ctypes.byref(c), None, ctypes.byref(h)) == -2
    lib.dgsem_destroy(h)


@pytest.mark.parametrize("dim,N,changes,rc,word", [
    (2, 3, dict(polydeg=0), -1, b"polydeg"),
    (2, 3, dict(polydeg=8), -1, b"polydeg"),
    (3, 6, {}, -2, b"N=6"),
    (3, 1, {}, -2, b"3D N=1"),
    (2, 3, dict(gamma=1.0), -1, b"gamma"),
    (2, 3, dict(equation=7), -1, b"equation"),
    (2, 3, dict(volume_integral=5), -1, b"volume integral"),
    (2, 3, dict(volume_flux=1), -1, b"volume flux"),
    (2, 3, dict(surface_flux=9), -1, b"surface or FV flux"),
    (2, 3, dict(fv_flux=-1), -1, b"surface or FV flux"),
    (2, 3, dict(geometry=4), -1, b"geometry"),
    (3, 3, dict(surface_flux=2), -2, b"CentralFlux"),
    (3, 3, dict(volume_flux=2), -2, b"CentralFlux"),
    (3, 3, dict(volume_integral=1, fv_flux=2), -2, b"CentralFlux"),
    (2, 7, dict(surface_flux=2), -2, b"CentralFlux"),
    (2, 3, dict(equation=1, mu=-1.0), -1, b"Navier-Stokes"),
])
def test_configuration_errors_without_gpu(dim, N, changes, rc, word):
    """Every configuration check of dgsem_create runs before device work, so
    it answers here with the error code and a message that names the field."""
    lib = _abi.load()
    c = _abi.Config()
    lib.dgsem_config_default(ctypes.byref(c), dim, N)
    for k, v in changes.items():
        setattr(c, k, v)
    h = ctypes.c_void_p()
    assert lib.dgsem_create(ctypes.byref(c), None, ctypes.byref(h)) == rc
    assert word in lib.dgsem_last_error(h), lib.dgsem_last_error(h)
    lib.dgsem_destroy(h)
