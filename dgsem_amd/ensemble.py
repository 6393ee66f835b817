"""Seeded ensembles of 2D Kelvin-Helmholtz cases for the batched path.

Case parameters come from SplitMix64 (Steele, Lea and Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014), a counter-based generator, so
the same seed gives the same cases on every machine:

* element grid (nx, ny), each drawn uniformly from {8, 16, 32, 64};
* polynomial degree N drawn uniformly from {2, 3, 4, 5};
* 8 Fourier modes per case, integer wavevector (kx, ky) in [1, 4]^2, phase in
  [0, 2 pi), amplitude in [0, 0.02].

The state is the Kelvin-Helmholtz initial condition of the reference's
example (examples/Euler2D/kelvin_helmholtz_hg/kelvin_helmholtz_hg.cpp:18-30)
on [-1, 1]^2 with the velocity perturbation
    du = sum_m a_m sin(pi (kx_m x + ky_m y) + phi_m),
    dv = sum_m a_m cos(pi (kx_m x + ky_m y) + phi_m),
which is periodic on the box.  The time step of each case follows the 2D rule
of the reference's examples (blast_wave_hg.cpp:113-114) with CFL 0.5:
dt = 0.5 * dx / ((2N + 1) * lambda_max), lambda_max = max over nodes and
directions of |v_d| + c of the initial state.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def choice(self, values):
        return values[(self.next_u64() >> 11) % len(values)]


@dataclass
class KhiCase:
    cells: tuple
    polydeg: int
    modes: np.ndarray  # [8][kx, ky, phase, amplitude]


def khi_ensemble(seed: int, ncases: int, grids=(8, 16, 32, 64), degrees=(2, 3, 4, 5),
                 nmodes: int = 8) -> list[KhiCase]:
    rng = SplitMix64(seed)
    cases = []
    for _ in range(ncases):
        nx, ny = rng.choice(grids), rng.choice(grids)
        N = rng.choice(degrees)
        modes = np.zeros((nmodes, 4))
        for m in range(nmodes):
            modes[m, 0] = rng.choice((1, 2, 3, 4))
            modes[m, 1] = rng.choice((1, 2, 3, 4))
            modes[m, 2] = 2.0 * np.pi * rng.uniform()
            modes[m, 3] = 0.02 * rng.uniform()
        cases.append(KhiCase((nx, ny), N, modes))
    return cases


def khi_state(coords: np.ndarray, modes: np.ndarray | None = None, gamma: float = 1.4):
    """Conservative state u(e, dof, var) at node coordinates [E][nn][2]."""
    x, y = coords[..., 0], coords[..., 1]
    B = np.tanh(15.0 * y + 7.5) - np.tanh(15.0 * y - 7.5)
    rho = 0.5 + 0.75 * B
    vx = 0.5 * (B - 1.0)
    vy = 0.1 * np.sin(2.0 * np.pi * x)
    if modes is not None:
        for kx, ky, phi, a in modes:
            arg = np.pi * (kx * x + ky * y) + phi
            vx = vx + a * np.sin(arg)
            vy = vy + a * np.cos(arg)
    p = np.ones_like(rho)
    u = np.empty(coords.shape[:-1] + (4,))
    u[..., 0] = rho
    u[..., 1] = rho * vx
    u[..., 2] = rho * vy
    u[..., 3] = p / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy)
    return u


def khi_dt(u: np.ndarray, polydeg: int, cells, gamma: float = 1.4, cfl: float = 0.5) -> float:
    rho = u[..., 0]
    vx, vy = u[..., 1] / rho, u[..., 2] / rho
    p = (gamma - 1.0) * (u[..., 3] - 0.5 * rho * (vx * vx + vy * vy))
    c = np.sqrt(gamma * p / rho)
    lam = max(float(np.max(np.abs(vx) + c)), float(np.max(np.abs(vy) + c)))
    dx = min(2.0 / cells[0], 2.0 / cells[1])
    return cfl * dx / ((2 * polydeg + 1) * lam)


def khi_configs(cases: list[KhiCase], geometry=None):
    """Solver configurations of the ensemble: the Kelvin-Helmholtz setup of
    the reference example (shock capturing, alpha_max 0.002, LLF fluxes)."""
    from . import _abi
    from .solver import SolverConfig
    geometry = _abi.GEOMETRY_NODAL if geometry is None else geometry
    return [SolverConfig(dim=2, polydeg=c.polydeg, cells=c.cells, domain_left=[-1.0, -1.0],
                         domain_right=[1.0, 1.0], volume_integral=_abi.SHOCK_CAPTURING,
                         alpha_max=0.002, alpha_min=0.001, geometry=geometry) for c in cases]


def khi_batch(cases: list[KhiCase], geometry=None):
    """(Batch, initial state arena, per-case dt) for the ensemble."""
    import torch

    from .solver import Batch
    batch = Batch(khi_configs(cases, geometry))
    u0 = batch.empty()
    dts = []
    for i, c in enumerate(cases):
        s = batch.case(i)
        ref = khi_state(s.node_coordinates(), c.modes)
        batch.view(u0, i).copy_(s.from_reference_layout(ref))
        dts.append(khi_dt(ref, c.polydeg, c.cells))
    batch.set_dt(dts)
    torch.cuda.synchronize()
    return batch, u0, dts
