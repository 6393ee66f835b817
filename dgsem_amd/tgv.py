"""Taylor-Green vortex initial state at Mach 0.1 (the reference's
examples/Euler3D/inviscid_tgv/inviscid_tgv.cpp:18-35 and
examples/NavierStokes3D/taylor_green_vortex/taylor_green_vortex.cpp), and
the 3D time-step rule of its driver (inviscid_tgv.cpp:116-117)."""
from __future__ import annotations

import numpy as np


def tgv_primitives(coords: np.ndarray, gamma: float = 1.4):
    """(rho, v[..., 3], p) at node coordinates [..., 3]."""
    x, y, z = coords[..., 0], coords[..., 1], coords[..., 2]
    rho = np.ones_like(x)
    v = np.stack([np.sin(x) * np.cos(y) * np.cos(z), -np.cos(x) * np.sin(y) * np.cos(z),
                  np.zeros_like(x)], axis=-1)
    p = 100.0 / gamma + (1.0 / 16.0) * (np.cos(2 * x) * np.cos(2 * z) + 2 * np.cos(2 * y)
                                       + 2 * np.cos(2 * x) + np.cos(2 * y) * np.cos(2 * z))
    return rho, v, p


def tgv_state(coords: np.ndarray, gamma: float = 1.4) -> np.ndarray:
    """Conservative state in the reference layout u[e][dof][var]."""
    rho, v, p = tgv_primitives(coords, gamma)
    E = p / (gamma - 1.0) + 0.5 * rho * np.sum(v * v, axis=-1)
    return np.concatenate([rho[..., None], rho[..., None] * v, E[..., None]], axis=-1)


def tgv_dt(cells: int, N: int, cfl: float = 0.5, gamma: float = 1.4) -> float:
    """cfl dx / (lambda_max 3 (2N + 1)) with lambda_max = max|v| + c of the
    initial state (|v| <= 1, c <= sqrt(gamma (100/gamma + 3/8)))."""
    dx = 2.0 * np.pi / cells
    lam = 1.0 + np.sqrt(gamma * (100.0 / gamma + 6.0 / 16.0))
    return cfl * dx / (lam * 3 * (2 * N + 1))
