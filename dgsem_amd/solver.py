"""Python mirror of the reference's solver interface over the HIP C ABI.

StructuredSolver mirrors DGSEM::StructuredSolver (structured_solver.hpp:19-125):
it is built from the equation, basis degree, volume/surface flux policies and
the mesh, and exposes calc_rhs, set_indicator_parameters and get_ndofs.
SSPRK3 mirrors DGSEM::SSPRK3 (SSPRK3.hpp:145-183).  Device memory and streams
come from PyTorch; all arithmetic runs in the HIP kernels of libdgsem_amd.so.
There is no CPU fallback: without a GPU or without the library these raise.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _abi

try:  # torch provides device memory and the stream
    import torch
except ImportError:  # pragma: no cover
    torch = None


@dataclass
class SolverConfig:
    dim: int
    polydeg: int
    cells: Sequence[int]
    domain_left: Sequence[float]
    domain_right: Sequence[float]
    equation: int = _abi.EULER
    gamma: float = 1.4
    mu: float | None = None
    prandtl: float | None = None
    volume_integral: int = _abi.SPLIT_FORM
    volume_flux: int = _abi.FLUX_CHANDRASHEKAR
    surface_flux: int = _abi.FLUX_LLF
    fv_flux: int = _abi.FLUX_LLF
    alpha_max: float = 0.5
    alpha_min: float = 0.001
    alpha_smooth: bool = False
    positivity_floor: float = 0.0
    geometry: int = _abi.GEOMETRY_NODAL
    # per-face boundary conditions (-x, +x, -y, +y[, -z, +z]); None: periodic
    face_boundary: Sequence[int] | None = None
    # constant Dirichlet states per face (conservative variables)
    face_state: Sequence[Sequence[float]] | None = None

    def to_c(self) -> _abi.Config:
        L = _abi.load()
        c = _abi.Config()
        L.dgsem_config_default(C.byref(c), self.dim, self.polydeg)
        for k in range(self.dim):
            c.cells[k] = int(self.cells[k])
            c.domain_left[k] = float(self.domain_left[k])
            c.domain_right[k] = float(self.domain_right[k])
        c.equation = self.equation
        c.gamma = self.gamma
        if self.mu is not None:
            c.mu = self.mu
        if self.prandtl is not None:
            c.prandtl = self.prandtl
        c.volume_integral = self.volume_integral
        c.volume_flux = self.volume_flux
        c.surface_flux = self.surface_flux
        c.fv_flux = self.fv_flux
        c.alpha_max = self.alpha_max
        c.alpha_min = self.alpha_min
        c.alpha_smooth = int(self.alpha_smooth)
        c.positivity_floor = self.positivity_floor
        c.geometry = self.geometry
        if self.face_boundary is not None:
            for f, kind in enumerate(self.face_boundary):
                c.face_boundary[f] = int(kind)
        if self.face_state is not None:
            for f, st in enumerate(self.face_state):
                if st is not None:
                    for v, x in enumerate(st):
                        c.face_state[f][v] = float(x)
        return c


def _ptr(t) -> int:
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError("expected a contiguous float64 tensor on the GPU")
    return t.data_ptr()


class StructuredSolver:
    """DGSEM right-hand side on one GPU (internal state layout [E][var][dof])."""

    def __init__(self, cfg: SolverConfig, stream=None, _handle=None, _owner=None):
        if torch is None or not torch.cuda.is_available():
            raise _abi.DgsemError("a ROCm GPU is required")
        self.cfg = cfg
        self._L = _abi.load()
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        self._owner = _owner  # a Batch that owns the handle, if any
        if _handle is not None:
            h = _handle
        else:
            h = C.c_void_p()
            rc = self._L.dgsem_create(C.byref(cfg.to_c()), C.c_void_p(self.stream.cuda_stream),
                                      C.byref(h))
            self._h = h
            if rc != 0:
                msg = self._L.dgsem_last_error(h).decode() if h.value else "create failed"
                if h.value:
                    self._L.dgsem_destroy(h)
                    self._h = C.c_void_p()
                raise _abi.DgsemError(f"dgsem_create: {msg} ({rc})")
        self._h = h
        E, nn, nv = C.c_int64(), C.c_int(), C.c_int()
        self._check(self._L.dgsem_sizes(h, C.byref(E), C.byref(nn), C.byref(nv)))
        self.n_elements, self.ndofs, self.nvars = E.value, nn.value, nv.value
        self.shape = (self.n_elements, self.nvars, self.ndofs)
        self.ref_shape = (self.n_elements, self.ndofs, self.nvars)

    def __del__(self):
        if getattr(self, "_owner", None) is not None:
            return  # the batch destroys its case handles
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.dgsem_destroy(self._h)
            self._h = C.c_void_p()

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise _abi.DgsemError(f"{self._L.dgsem_last_error(self._h).decode()} ({rc})")

    # -- reference-interface methods ---------------------------------------
    def get_ndofs(self) -> int:
        return self.ndofs

    def set_indicator_parameters(self, alpha_max: float, alpha_min: float, alpha_smooth: bool) -> None:
        self._check(self._L.dgsem_set_indicator_parameters(self._h, alpha_max, alpha_min, int(alpha_smooth)))

    def set_boundary_values(self, face: int, values) -> None:
        """Per-node Dirichlet states of a face (device tensor [nodes][nvars],
        kept referenced here), or None for the face's constant state."""
        if not hasattr(self, "_bc_values"):
            self._bc_values = {}
        self._bc_values[face] = values
        ptr = None if values is None else _ptr(values)
        self._check(self._L.dgsem_set_boundary_values(self._h, int(face), ptr))

    def calc_rhs(self, u, du, t: float = 0.0) -> None:
        self._check(self._L.dgsem_rhs(self._h, _ptr(u), _ptr(du), float(t)))

    # -- extras --------------------------------------------------------------
    def empty(self):
        return torch.empty(self.shape, dtype=torch.float64, device="cuda")

    def node_coordinates(self) -> np.ndarray:
        out = np.zeros((self.n_elements, self.ndofs, self.cfg.dim))
        self._check(self._L.dgsem_node_coordinates(self._h, out.ctypes.data))
        return out

    def reference_coordinates(self) -> np.ndarray:
        """Reference-cube position of every node, [E][nn][dim] (the argument
        of a mapping)."""
        out = np.zeros((self.n_elements, self.ndofs, self.cfg.dim))
        self._check(self._L.dgsem_reference_coordinates(self._h, out.ctypes.data))
        return out

    def set_node_coordinates(self, coords) -> None:
        """Mapped node coordinates [E][nn][dim] (numpy or a GPU tensor); the
        geometry is recomputed on the GPU."""
        if not torch.is_tensor(coords):
            coords = torch.from_numpy(np.ascontiguousarray(coords, dtype=np.float64)).cuda()
        self._check(self._L.dgsem_set_node_coordinates(self._h, _ptr(coords)))

    def apply_mapping(self, mapping) -> None:
        """Curvilinear mesh from a vectorised mapping f(ref [..., dim]) ->
        physical coordinates [..., dim], evaluated at the reference
        positions of the nodes (StructuredElementInitializer with a user
        Mapping)."""
        self.set_node_coordinates(np.asarray(mapping(self.reference_coordinates()), dtype=np.float64))

    def contravariant_vectors(self) -> np.ndarray:
        out = np.zeros((self.n_elements, self.ndofs, self.cfg.dim, self.cfg.dim))
        self._check(self._L.dgsem_contravariant_vectors(self._h, out.ctypes.data))
        return out

    def inverse_jacobian(self) -> np.ndarray:
        out = np.zeros((self.n_elements, self.ndofs))
        self._check(self._L.dgsem_inverse_jacobian(self._h, out.ctypes.data))
        return out

    def from_reference_layout(self, u_ref):
        """u(e, dof, var) (numpy or tensor) -> internal [e][var][dof] on the GPU."""
        src = torch.as_tensor(np.ascontiguousarray(u_ref) if isinstance(u_ref, np.ndarray) else u_ref,
                              dtype=torch.float64).to("cuda").contiguous()
        dst = self.empty()
        self._check(self._L.dgsem_pack_state(self._h, _ptr(src), _ptr(dst)))
        return dst

    def to_reference_layout(self, u) -> np.ndarray:
        dst = torch.empty(self.ref_shape, dtype=torch.float64, device="cuda")
        self._check(self._L.dgsem_unpack_state(self._h, _ptr(u), _ptr(dst)))
        self.synchronize()
        return dst.cpu().numpy()

    def pack_layout(self, flat, layout: int):
        """Flat device buffer of a reference view in `layout` (LAYOUT_RIGHT or
        LAYOUT_LEFT) -> internal [e][var][dof]."""
        dst = self.empty()
        self._check(self._L.dgsem_pack_state_layout(self._h, int(layout), _ptr(flat), _ptr(dst)))
        return dst

    def unpack_layout(self, u, layout: int):
        """Internal state -> flat device buffer of a reference view in `layout`."""
        dst = torch.empty(u.numel(), dtype=torch.float64, device="cuda")
        self._check(self._L.dgsem_unpack_state_layout(self._h, int(layout), _ptr(u), _ptr(dst)))
        return dst

    def alpha(self):
        a = torch.empty(self.n_elements, dtype=torch.float64, device="cuda")
        self._check(self._L.dgsem_get_alpha(self._h, _ptr(a)))
        return a

    def stage_alpha(self, stage: int):
        """Blending factor the last SSPRK3 stage of that number (1-3)
        recorded, or the last calc_rhs (0)."""
        a = torch.empty(self.n_elements, dtype=torch.float64, device="cuda")
        self._check(self._L.dgsem_get_stage_alpha(self._h, int(stage), _ptr(a)))
        return a

    def freeze_alpha(self, alpha) -> None:
        """Blend with this per-element alpha (device tensor, kept referenced
        here) instead of the indicator's, in the forward evaluations and as
        the frozen alpha of the derivatives; None returns to the indicator."""
        self._frozen_alpha = None if alpha is None else alpha.contiguous()
        ptr = None if alpha is None else _ptr(self._frozen_alpha)
        self._check(self._L.dgsem_freeze_alpha(self._h, ptr))

    def ssprk3_stage(self, stage: int, dt: float, u_base, u_k, u_out) -> None:
        base = _ptr(u_base) if u_base is not None else None
        self._check(self._L.dgsem_ssprk3_stage(self._h, stage, float(dt), base, _ptr(u_k), _ptr(u_out)))

    def ssprk3_steps(self, u, dt: float, nsteps: int, tmp1, tmp2) -> None:
        self._check(self._L.dgsem_ssprk3_steps(self._h, _ptr(u), float(dt), int(nsteps), _ptr(tmp1), _ptr(tmp2)))

    # -- tangent / adjoint (no reference counterpart) ------------------------
    def rhs_jvp(self, u, du, out=None):
        """out = J(u) du, J = dL/du."""
        out = self.empty() if out is None else out
        self._check(self._L.dgsem_rhs_jvp(self._h, _ptr(u), _ptr(du), _ptr(out)))
        return out

    def rhs_vjp(self, u, w, out=None):
        """out = J(u)^T w."""
        out = self.empty() if out is None else out
        self._check(self._L.dgsem_rhs_vjp(self._h, _ptr(u), _ptr(w), _ptr(out)))
        return out

    def ssprk3_stages(self, u0, dt: float, u1=None, u2=None):
        """Stage states u1, u2 of one SSPRK3 step from u0."""
        u1 = self.empty() if u1 is None else u1
        u2 = self.empty() if u2 is None else u2
        self.ssprk3_stage(1, dt, None, u0, u1)
        self.ssprk3_stage(2, dt, u0, u1, u2)
        return u1, u2

    def step_jvp(self, dt: float, u0, u1, u2, du0, out=None, tmp1=None, tmp2=None):
        out = self.empty() if out is None else out
        tmp1 = self.empty() if tmp1 is None else tmp1
        tmp2 = self.empty() if tmp2 is None else tmp2
        self._check(self._L.dgsem_ssprk3_step_jvp(self._h, float(dt), _ptr(u0), _ptr(u1), _ptr(u2),
                                                   _ptr(du0), _ptr(out), _ptr(tmp1), _ptr(tmp2)))
        return out

    def step_vjp(self, dt: float, u0, u1, u2, ubar, out=None, tmp1=None, tmp2=None):
        out = self.empty() if out is None else out
        tmp1 = self.empty() if tmp1 is None else tmp1
        tmp2 = self.empty() if tmp2 is None else tmp2
        self._check(self._L.dgsem_ssprk3_step_vjp(self._h, float(dt), _ptr(u0), _ptr(u1), _ptr(u2),
                                                   _ptr(ubar), _ptr(out), _ptr(tmp1), _ptr(tmp2)))
        return out

    def kinetic_energy(self, u, result=None):
        """Kinetic-energy quadrature of u into a one-element device tensor."""
        result = torch.empty(1, dtype=torch.float64, device="cuda") if result is None else result
        self._check(self._L.dgsem_kinetic_energy(self._h, _ptr(u), _ptr(result)))
        return result

    def kinetic_energy_grad(self, u, scale: float, grad) -> None:
        """grad += scale * d(kinetic energy)/du."""
        self._check(self._L.dgsem_kinetic_energy_grad(self._h, _ptr(u), float(scale), _ptr(grad)))

    def analyze(self, u) -> dict:
        """One-pass analysis of u (synchronises): volume, kinetic-energy and
        entropy integrals, their volume means, the number of non-finite
        nodes and the largest signal speed max_d |v_d| + c."""
        out = torch.empty(5, dtype=torch.float64, device="cuda")
        self._check(self._L.dgsem_analyze(self._h, _ptr(u), _ptr(out)))
        vol, ke, ent, bad, vmax = out.cpu().tolist()
        return {"volume": vol, "kinetic_energy": ke, "entropy": ent,
                "mean_kinetic_energy": ke / vol if vol else float("nan"),
                "mean_entropy": ent / vol if vol else float("nan"),
                "nonfinite_nodes": int(bad), "max_speed": vmax}

    def synchronize(self) -> None:
        self._check(self._L.dgsem_synchronize(self._h))

    def write_vtu(self, u, path: str) -> None:
        """VTK XML file of Lagrange cells holding u (dgsem_write_vtu)."""
        self._check(self._L.dgsem_write_vtu(self._h, _ptr(u), str(path).encode()))


class SSPRK3:
    """Three-stage SSP Runge-Kutta in Shu-Osher form on a solver's state,
    SSPRK3.hpp:145-183; each stage is one fused kernel."""

    def __init__(self, solver: StructuredSolver, final_time: float = 0.0):
        self.solver = solver
        self.tmp1 = solver.empty()
        self.tmp2 = solver.empty()
        self.time = 0.0
        self.final_time = final_time
        self.iter = 0
        self.observers = []
        self.has_nan = False

    def add_observer(self, observer) -> None:
        self.observers.append(observer)

    def step(self, u, dt: float) -> None:
        s = self.solver
        s.ssprk3_stage(1, dt, None, u, self.tmp1)
        s.ssprk3_stage(2, dt, u, self.tmp1, self.tmp2)
        s.ssprk3_stage(3, dt, u, self.tmp2, u)

    def solve(self, u, dt: float) -> None:
        """TimeIntegrator::solve (integrator_base.hpp:21-40): fixed dt while
        t < t_final; observers run before each step and may stop the run."""
        while self.time < self.final_time:
            for obs in self.observers:
                self.has_nan = bool(obs.on_step(self.iter, self.time)) or self.has_nan
            self.step(u, dt)
            self.time += dt
            self.iter += 1
            if self.has_nan:
                break
        for obs in self.observers:
            obs.on_finish(self.iter, self.time)


class Batch:
    """Ensemble of independent 2D Euler cases advanced together
    (dgsem_batch_* in include/dgsem_amd.h).  States live in arenas of
    `total` doubles; case i is the slice `view(arena, i)` in the internal
    layout of `case(i)`, a StructuredSolver over the case's own handle."""

    def __init__(self, cfgs: Sequence[SolverConfig], stream=None):
        if torch is None or not torch.cuda.is_available():
            raise _abi.DgsemError("a ROCm GPU is required")
        self._L = _abi.load()
        self.cfgs = list(cfgs)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        arr = (_abi.Config * len(self.cfgs))(*[c.to_c() for c in self.cfgs])
        b = C.c_void_p()
        rc = self._L.dgsem_batch_create(arr, len(self.cfgs), C.c_void_p(self.stream.cuda_stream),
                                        C.byref(b))
        self._b = b
        if rc != 0:
            msg = self._L.dgsem_batch_last_error(b).decode() if b.value else "create failed"
            if b.value:
                self._L.dgsem_batch_destroy(b)
                self._b = C.c_void_p()
            raise _abi.DgsemError(f"dgsem_batch_create: {msg} ({rc})")
        n, total = C.c_int(), C.c_int64()
        self._check(self._L.dgsem_batch_sizes(b, C.byref(n), C.byref(total)))
        self.ncases, self.total = n.value, total.value
        self.offsets, self.counts, self._cases = [], [], []
        for i in range(self.ncases):
            off, cnt, h = C.c_int64(), C.c_int64(), C.c_void_p()
            self._check(self._L.dgsem_batch_case(b, i, C.byref(off), C.byref(cnt), C.byref(h)))
            self.offsets.append(off.value)
            self.counts.append(cnt.value)
            self._cases.append(StructuredSolver(self.cfgs[i], self.stream, _handle=h, _owner=self))

    def __del__(self):
        if getattr(self, "_b", None) is not None and self._b.value:
            self._cases = []
            self._L.dgsem_batch_destroy(self._b)
            self._b = C.c_void_p()

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise _abi.DgsemError(f"{self._L.dgsem_batch_last_error(self._b).decode()} ({rc})")

    def case(self, i: int) -> StructuredSolver:
        return self._cases[i]

    def empty(self):
        return torch.empty(self.total, dtype=torch.float64, device="cuda")

    def view(self, arena, i: int):
        """Case i's slice of an arena, shaped like the case's internal state."""
        o, c = self.offsets[i], self.counts[i]
        return arena[o:o + c].view(self._cases[i].shape)

    def set_dt(self, dt: Sequence[float]) -> None:
        a = np.ascontiguousarray(dt, dtype=np.float64)
        if a.shape != (self.ncases,):
            raise ValueError("one dt per case")
        self._check(self._L.dgsem_batch_set_dt(self._b, a.ctypes.data_as(C.POINTER(C.c_double))))

    def rhs(self, u, du) -> None:
        self._check(self._L.dgsem_batch_rhs(self._b, _ptr(u), _ptr(du)))

    def ssprk3_steps(self, u, nsteps: int, tmp1, tmp2) -> None:
        self._check(self._L.dgsem_batch_ssprk3_steps(self._b, _ptr(u), int(nsteps), _ptr(tmp1),
                                                     _ptr(tmp2)))

    def tuning(self) -> dict:
        blk, grid = C.c_int(), C.c_int()
        tiles = (C.c_int * 8)()
        self._check(self._L.dgsem_batch_get_tuning(self._b, C.byref(blk), tiles, C.byref(grid)))
        return {"block": blk.value, "tile_elems": list(tiles), "grid": grid.value}

    def set_tuning(self, block: int, tile_elems: Sequence[int] | None = None, grid: int = 0) -> None:
        tiles = (C.c_int * 8)(*(list(tile_elems) if tile_elems is not None else [0] * 8))
        self._check(self._L.dgsem_batch_set_tuning(self._b, int(block), tiles, int(grid)))

    def autotune(self, u, scratch, reps: int = 5) -> dict:
        self._check(self._L.dgsem_batch_autotune(self._b, _ptr(u), _ptr(scratch), int(reps)))
        return self.tuning()

