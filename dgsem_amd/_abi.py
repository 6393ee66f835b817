"""ctypes binding of libdgsem_amd.so (declarations in include/dgsem_amd.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdgsem_amd.so")

EULER, NAVIER_STOKES = 0, 1
SPLIT_FORM, SHOCK_CAPTURING = 0, 1
FLUX_CHANDRASHEKAR = 0
FLUX_LLF, FLUX_HLLC = 0, 1
FLUX_CENTRAL = 2  # volume, surface or subcell flux (2D)
GEOMETRY_NODAL, GEOMETRY_AFFINE = 0, 1
BC_PERIODIC, BC_SLIP_WALL, BC_OUTFLOW, BC_DIRICHLET = 0, 1, 2, 3
BC_NS_NOSLIP_WALL, BC_NS_SLIP_WALL, BC_MIXED_DIRICHLET_SLIP_WALL = 4, 5, 6
LAYOUT_RIGHT, LAYOUT_LEFT = 0, 1

EXPORTED = (
    "dgsem_abi_version", "dgsem_config_default", "dgsem_create", "dgsem_destroy",
    "dgsem_last_error", "dgsem_sizes", "dgsem_set_indicator_parameters",
    "dgsem_node_coordinates", "dgsem_contravariant_vectors", "dgsem_inverse_jacobian",
    "dgsem_set_boundary_values", "dgsem_write_vtu", "dgsem_reference_coordinates",
    "dgsem_set_node_coordinates", "dgsem_pack_state", "dgsem_unpack_state", "dgsem_pack_state_layout",
    "dgsem_unpack_state_layout", "dgsem_rhs", "dgsem_get_alpha", "dgsem_get_stage_alpha",
    "dgsem_freeze_alpha",
    "dgsem_ssprk3_stage", "dgsem_ssprk3_steps", "dgsem_synchronize",
    "dgsem_batch_create", "dgsem_batch_destroy", "dgsem_batch_last_error", "dgsem_batch_sizes",
    "dgsem_batch_case", "dgsem_batch_set_dt", "dgsem_batch_rhs", "dgsem_batch_ssprk3_steps",
    "dgsem_batch_get_tuning", "dgsem_batch_set_tuning", "dgsem_batch_autotune",
    "dgsem_rhs_jvp", "dgsem_rhs_vjp", "dgsem_ssprk3_step_jvp", "dgsem_ssprk3_step_vjp",
    "dgsem_kinetic_energy", "dgsem_kinetic_energy_grad", "dgsem_analyze",
    "dgsem_analyze_noalloc",
)


class Config(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("polydeg", C.c_int), ("cells", C.c_int * 3),
        ("domain_left", C.c_double * 3), ("domain_right", C.c_double * 3),
        ("boundary", C.c_int * 3), ("equation", C.c_int), ("gamma", C.c_double),
        ("mu", C.c_double), ("prandtl", C.c_double), ("volume_integral", C.c_int),
        ("volume_flux", C.c_int), ("surface_flux", C.c_int), ("fv_flux", C.c_int),
        ("alpha_max", C.c_double), ("alpha_min", C.c_double), ("alpha_smooth", C.c_int),
        ("positivity_floor", C.c_double), ("geometry", C.c_int),
        ("face_boundary", C.c_int * 6), ("face_state", (C.c_double * 5) * 6),
    ]


class DgsemError(RuntimeError):
    pass


_lib = None


def load(path: str | None = None):
    """Loads the HIP library; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("DGSEM_AMD_LIBRARY") or LIB_PATH
    if not os.path.exists(path):
        raise DgsemError(f"{path} not built; run __graft_entry__.build() or dgsem_amd/build.py")
    L = C.CDLL(path)
    vp, dp, i64p, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int)
    L.dgsem_abi_version.restype = C.c_int
    L.dgsem_config_default.argtypes = [C.POINTER(Config), C.c_int, C.c_int]
    L.dgsem_config_default.restype = None
    L.dgsem_create.argtypes = [C.POINTER(Config), vp, C.POINTER(vp)]
    L.dgsem_destroy.argtypes = [vp]
    L.dgsem_last_error.argtypes = [vp]
    L.dgsem_last_error.restype = C.c_char_p
    L.dgsem_sizes.argtypes = [vp, i64p, ip, ip]
    L.dgsem_set_indicator_parameters.argtypes = [vp, C.c_double, C.c_double, C.c_int]
    for name in ("dgsem_node_coordinates", "dgsem_contravariant_vectors", "dgsem_inverse_jacobian"):
        getattr(L, name).argtypes = [vp, vp]
    L.dgsem_pack_state.argtypes = [vp, vp, vp]
    L.dgsem_set_boundary_values.argtypes = [vp, C.c_int, vp]
    L.dgsem_write_vtu.argtypes = [vp, vp, C.c_char_p]
    L.dgsem_reference_coordinates.argtypes = [vp, vp]
    L.dgsem_set_node_coordinates.argtypes = [vp, vp]
    L.dgsem_unpack_state.argtypes = [vp, vp, vp]
    L.dgsem_pack_state_layout.argtypes = [vp, C.c_int, vp, vp]
    L.dgsem_unpack_state_layout.argtypes = [vp, C.c_int, vp, vp]
    L.dgsem_rhs.argtypes = [vp, vp, vp, C.c_double]
    L.dgsem_get_alpha.argtypes = [vp, vp]
    L.dgsem_get_stage_alpha.argtypes = [vp, C.c_int, vp]
    L.dgsem_freeze_alpha.argtypes = [vp, vp]
    L.dgsem_ssprk3_stage.argtypes = [vp, C.c_int, C.c_double, vp, vp, vp]
    L.dgsem_ssprk3_steps.argtypes = [vp, vp, C.c_double, C.c_int, vp, vp]
    L.dgsem_synchronize.argtypes = [vp]
    L.dgsem_batch_create.argtypes = [C.POINTER(Config), C.c_int, vp, C.POINTER(vp)]
    L.dgsem_batch_destroy.argtypes = [vp]
    L.dgsem_batch_last_error.argtypes = [vp]
    L.dgsem_batch_last_error.restype = C.c_char_p
    L.dgsem_batch_sizes.argtypes = [vp, ip, i64p]
    L.dgsem_batch_case.argtypes = [vp, C.c_int, i64p, i64p, C.POINTER(vp)]
    L.dgsem_batch_set_dt.argtypes = [vp, dp]
    L.dgsem_batch_rhs.argtypes = [vp, vp, vp]
    L.dgsem_batch_ssprk3_steps.argtypes = [vp, vp, C.c_int, vp, vp]
    L.dgsem_batch_get_tuning.argtypes = [vp, ip, ip, ip]
    L.dgsem_batch_set_tuning.argtypes = [vp, C.c_int, ip, C.c_int]
    L.dgsem_batch_autotune.argtypes = [vp, vp, vp, C.c_int]
    L.dgsem_rhs_jvp.argtypes = [vp, vp, vp, vp]
    L.dgsem_rhs_vjp.argtypes = [vp, vp, vp, vp]
    L.dgsem_ssprk3_step_jvp.argtypes = [vp, C.c_double, vp, vp, vp, vp, vp, vp, vp]
    L.dgsem_ssprk3_step_vjp.argtypes = [vp, C.c_double, vp, vp, vp, vp, vp, vp, vp]
    L.dgsem_kinetic_energy.argtypes = [vp, vp, vp]
    L.dgsem_kinetic_energy_grad.argtypes = [vp, vp, C.c_double, vp]
    L.dgsem_analyze.argtypes = [vp, vp, vp]
    L.dgsem_analyze_noalloc.argtypes = [vp, vp, vp]
    for name in EXPORTED:
        if name not in ("dgsem_config_default", "dgsem_last_error", "dgsem_batch_last_error"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L
