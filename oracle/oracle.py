"""ctypes wrapper around the CPU restatement in dgsem_oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product package
(dgsem_amd) never imports this module.

Arrays use the reference's LayoutRight layouts
(u[ielem][jelem][kelem][dof][var], element index (ie*ny + je)*nz + ke).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libdgsem_oracle.so")
_lib = None

EULER, NAVIER_STOKES = 0, 1
SPLIT_FORM, SHOCK_CAPTURING = 0, 1
FLUX_LLF, FLUX_HLLC, FLUX_CENTRAL = 0, 1, 2
VOLUME_CHANDRASHEKAR, VOLUME_CENTRAL = 0, 2
BC_PERIODIC, BC_SLIP_WALL, BC_OUTFLOW, BC_DIRICHLET = 0, 1, 2, 3
BC_NS_NOSLIP_WALL, BC_NS_SLIP_WALL, BC_MIXED_DIRICHLET_SLIP_WALL = 4, 5, 6

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the restatement with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    L.orc_basis.argtypes = [C.c_int] + [_dp] * 9
    L.orc_create.restype = vp
    L.orc_create.argtypes = [C.c_int] * 5 + [_dp, _dp, np.ctypeslib.ndpointer(dtype=np.int32)]
    L.orc_destroy.argtypes = [vp]
    L.orc_set_equation.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double]
    L.orc_set_volume.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double]
    L.orc_set_surface_flux.argtypes = [vp, C.c_int]
    L.orc_set_volume_flux.argtypes = [vp, C.c_int]
    L.orc_set_positivity_floor.argtypes = [vp, C.c_double]
    L.orc_set_threads.argtypes = [vp, C.c_int]
    L.orc_num_elements.restype = C.c_int64
    L.orc_num_elements.argtypes = [vp]
    L.orc_nvars.argtypes = [vp]
    L.orc_ndofs.argtypes = [vp]
    for name in ("orc_get_coordinates", "orc_get_jacobian", "orc_get_contravariant",
                 "orc_get_inverse_jacobian"):
        getattr(L, name).argtypes = [vp, _dp]
    L.orc_get_left_neighbors.argtypes = [vp, _ip]
    L.orc_set_metrics.argtypes = [vp, _dp, _dp]
    L.orc_set_coordinates.argtypes = [vp, _dp]
    L.orc_init_khi.argtypes = [vp, _dp]
    L.orc_init_tgv.argtypes = [vp, _dp]
    L.orc_init_sedov.argtypes = [vp, _dp, C.c_double, C.c_double, C.c_double]
    L.orc_rhs.argtypes = [vp, _dp, _dp, vp]
    L.orc_volume_integral.argtypes = [vp, _dp, _dp, vp]
    L.orc_calc_alpha.argtypes = [vp, _dp, _dp]
    L.orc_interface_flux.argtypes = [vp, _dp, _dp]
    L.orc_surface_integral.argtypes = [vp, _dp, _dp]
    L.orc_apply_jacobian.argtypes = [vp, _dp]
    L.orc_gradients.argtypes = [vp, _dp, _dp]
    L.orc_ssprk3.argtypes = [vp, _dp, C.c_double, C.c_int]
    L.orc_apply_floor.argtypes = [vp, _dp]
    L.orc_set_alpha_override.argtypes = [vp, vp]
    L.orc_gradient_variables.argtypes = [vp, _dp, _dp]
    L.orc_viscous_flux.argtypes = [vp, _dp, _dp, C.c_int, _dp]
    L.orc_ln_mean.restype = C.c_double
    L.orc_ln_mean.argtypes = [C.c_double, C.c_double]
    for name in ("orc_flux_chandrashekar", "orc_flux_llf", "orc_flux_hllc", "orc_flux_central"):
        getattr(L, name).argtypes = [C.c_int, C.c_double, _dp, _dp, _dp, _dp]
    L.orc_physical_flux.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp]
    L.orc_kinetic_energy.restype = C.c_double
    L.orc_kinetic_energy.argtypes = [vp, _dp]
    L.orc_set_boundary.argtypes = [vp, C.c_int, C.c_int, vp, vp]
    L.orc_analyze.argtypes = [vp, _dp, _dp]
    _lib = L
    return L


def basis(N: int) -> dict:
    n = N + 1
    out = {k: np.zeros(n) for k in ("nodes", "weights", "inv_weights", "lhat_left", "lhat_right")}
    mats = {k: np.zeros(n * n) for k in ("D", "Dsplit", "Dhat", "Vinv")}
    lib().orc_basis(N, out["nodes"], out["weights"], out["inv_weights"], mats["D"],
                    mats["Dsplit"], mats["Dhat"], mats["Vinv"], out["lhat_left"],
                    out["lhat_right"])
    for k, v in mats.items():
        out[k] = v.reshape(n, n)
    return out


def ln_mean(x: float, y: float) -> float:
    return lib().orc_ln_mean(float(x), float(y))


def _flux(name, dim, gamma, ul, ur, n):
    f = np.zeros(dim + 2)
    getattr(lib(), name)(dim, gamma, np.ascontiguousarray(ul, dtype=np.float64),
                         np.ascontiguousarray(ur, dtype=np.float64),
                         np.ascontiguousarray(n, dtype=np.float64), f)
    return f


def flux_chandrashekar(dim, gamma, ul, ur, n):
    return _flux("orc_flux_chandrashekar", dim, gamma, ul, ur, n)


def flux_llf(dim, gamma, ul, ur, n):
    return _flux("orc_flux_llf", dim, gamma, ul, ur, n)


def flux_hllc(dim, gamma, ul, ur, n):
    return _flux("orc_flux_hllc", dim, gamma, ul, ur, n)


def flux_central(dim, gamma, ul, ur, n):
    return _flux("orc_flux_central", dim, gamma, ul, ur, n)


def physical_flux(dim, gamma, u, n):
    f = np.zeros(dim + 2)
    lib().orc_physical_flux(dim, gamma, np.ascontiguousarray(u, dtype=np.float64),
                            np.ascontiguousarray(n, dtype=np.float64), f)
    return f


class Case:
    """One mesh + solver configuration of the restatement."""

    def __init__(self, dim, N, cells, left, right, periodic=None, *, eq=EULER,
                 gamma=1.4, mu=None, prandtl=None, volume=SPLIT_FORM, fv_flux=FLUX_LLF,
                 alpha_max=0.5, alpha_min=0.001, surface_flux=FLUX_LLF,
                 positivity_floor=0.0, threads=0, volume_flux=VOLUME_CHANDRASHEKAR):
        self.dim, self.N = dim, N
        cells = list(cells) + [1] * (3 - len(cells))
        self.cells = tuple(cells[:dim])
        periodic = periodic or [1] * dim
        L = lib()
        self._h = L.orc_create(dim, N, cells[0], cells[1], cells[2],
                               np.asarray(left, dtype=np.float64),
                               np.asarray(right, dtype=np.float64),
                               np.asarray(periodic, dtype=np.int32))
        if not self._h:
            raise ValueError("orc_create failed")
        if mu is None:
            mu = 6.25e-4 if dim == 3 else 5.0e-3
        if prandtl is None:
            prandtl = 0.72 if dim == 3 else 0.73
        L.orc_set_equation(self._h, eq, gamma, mu, prandtl)
        L.orc_set_volume(self._h, volume, fv_flux, alpha_max, alpha_min)
        L.orc_set_surface_flux(self._h, surface_flux)
        L.orc_set_volume_flux(self._h, volume_flux)
        L.orc_set_positivity_floor(self._h, positivity_floor)
        L.orc_set_threads(self._h, threads)
        self.E = int(L.orc_num_elements(self._h))
        self.nvars = dim + 2
        self.ndofs = (N + 1) ** dim
        self.shape = (self.E, self.ndofs, self.nvars)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_destroy(self._h)
            self._h = None

    def set_threads(self, t: int) -> None:
        lib().orc_set_threads(self._h, t)

    def zeros(self):
        return np.zeros(self.shape)

    def coordinates(self):
        out = np.zeros((self.E, self.ndofs, self.dim))
        lib().orc_get_coordinates(self._h, out)
        return out

    def jacobian(self):
        out = np.zeros((self.E, self.ndofs, self.dim, self.dim))
        lib().orc_get_jacobian(self._h, out)
        return out

    def contravariant(self):
        out = np.zeros((self.E, self.ndofs, self.dim, self.dim))
        lib().orc_get_contravariant(self._h, out)
        return out

    def inverse_jacobian(self):
        out = np.zeros((self.E, self.ndofs))
        lib().orc_get_inverse_jacobian(self._h, out)
        return out

    def left_neighbors(self):
        out = np.zeros((self.E, self.dim, self.dim), dtype=np.int64)
        lib().orc_get_left_neighbors(self._h, out)
        return out

    def reference_coordinates(self):
        """Reference-cube position of every node before the mapping
        (calc_node_coordinates, data_container.hpp:193-221, :371-414)."""
        n = self.N + 1
        nodes = basis(self.N)["nodes"]
        cells = list(self.cells)
        out = np.zeros((self.E, self.ndofs, self.dim))
        dof = np.arange(self.ndofs)
        nd = [dof % n, (dof // n) % n, dof // (n * n)]
        e = np.arange(self.E)
        if self.dim == 2:
            idx = [e // cells[1], e % cells[1]]
        else:
            idx = [e // (cells[1] * cells[2]), (e // cells[2]) % cells[1], e % cells[2]]
        for k in range(self.dim):
            delta = 2.0 / cells[k]
            off = -1.0 + idx[k].astype(np.float64) * delta + 0.5 * delta
            out[..., k] = off[:, None] + 0.5 * delta * nodes[nd[k]][None, :]
        return out

    def set_coordinates(self, coords):
        """Node coordinates of any mapping; metrics are recomputed."""
        lib().orc_set_coordinates(self._h, np.ascontiguousarray(coords, dtype=np.float64))

    def set_metrics(self, Ja, invJ):
        lib().orc_set_metrics(self._h, np.ascontiguousarray(Ja, dtype=np.float64),
                              np.ascontiguousarray(invJ, dtype=np.float64))

    def init_khi(self):
        u = self.zeros()
        lib().orc_init_khi(self._h, u)
        return u

    def init_tgv(self):
        u = self.zeros()
        lib().orc_init_tgv(self._h, u)
        return u

    def init_sedov(self, E0=1.0, sigma=0.0625, p_background=1e-5):
        u = self.zeros()
        lib().orc_init_sedov(self._h, u, E0, sigma, p_background)
        return u

    def rhs(self, u, with_alpha=False):
        u = np.ascontiguousarray(u, dtype=np.float64)
        du = self.zeros()
        if with_alpha:
            alpha = np.zeros(self.E)
            lib().orc_rhs(self._h, u, du, alpha.ctypes.data_as(C.c_void_p))
            return du, alpha
        lib().orc_rhs(self._h, u, du, None)
        return du

    def alpha(self, u):
        a = np.zeros(self.E)
        lib().orc_calc_alpha(self._h, np.ascontiguousarray(u, dtype=np.float64), a)
        return a

    def volume_integral(self, u):
        du = self.zeros()
        lib().orc_volume_integral(self._h, np.ascontiguousarray(u, dtype=np.float64), du, None)
        return du

    def interface_flux(self, u):
        fn = (self.N + 1) ** (self.dim - 1)
        sf = np.zeros((self.E, 2 * self.dim * fn, self.nvars))
        lib().orc_interface_flux(self._h, np.ascontiguousarray(u, dtype=np.float64), sf)
        return sf

    def gradients(self, u):
        g = np.zeros((self.E, self.ndofs, self.nvars, self.dim))
        lib().orc_gradients(self._h, np.ascontiguousarray(u, dtype=np.float64), g)
        return g

    def ssprk3(self, u, dt, nsteps):
        u = np.array(u, dtype=np.float64, order="C", copy=True)
        lib().orc_ssprk3(self._h, u, float(dt), int(nsteps))
        return u

    def gradient_variables(self, u):
        q = np.zeros(self.nvars)
        lib().orc_gradient_variables(self._h, np.ascontiguousarray(u, dtype=np.float64), q)
        return q

    def viscous_flux(self, q, grad, dim):
        """grad: [dim][nvars] gradient of the gradient variables."""
        f = np.zeros(self.nvars)
        lib().orc_viscous_flux(self._h, np.ascontiguousarray(q, dtype=np.float64),
                               np.ascontiguousarray(grad, dtype=np.float64), int(dim), f)
        return f

    def set_alpha_override(self, alpha):
        """Blend every later shock-capturing RHS with this alpha [E] instead
        of the indicator's (None restores the indicator)."""
        self._alpha_keep = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        lib().orc_set_alpha_override(self._h, None if alpha is None else self._alpha_keep.ctypes.data)

    def apply_floor(self, u):
        """The positivity floor of a stage update on a copy of u."""
        u = np.array(u, dtype=np.float64, order="C", copy=True)
        lib().orc_apply_floor(self._h, u)
        return u

    def kinetic_energy(self, u):
        return lib().orc_kinetic_energy(self._h, np.ascontiguousarray(u, dtype=np.float64))

    def set_boundary(self, face: int, kind: int, state=None, values=None):
        """Boundary condition of one face of a 2D mesh (face = 2*dim + upper);
        values: per-node Dirichlet states [face nodes][nvars] (kept alive here)."""
        st = None
        if state is not None:  # the C side reads five values
            st = np.zeros(5)
            st[:len(state)] = np.asarray(state, dtype=np.float64)
        va = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        if not hasattr(self, "_bc_keep"):
            self._bc_keep = {}
        self._bc_keep[face] = (st, va)
        lib().orc_set_boundary(self._h, face, kind, None if st is None else st.ctypes.data,
                               None if va is None else va.ctypes.data)

    def analyze(self, u):
        out = np.zeros(5)
        lib().orc_analyze(self._h, np.ascontiguousarray(u, dtype=np.float64), out)
        return out
