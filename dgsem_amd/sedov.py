"""Gaussian Sedov blast initial state and its time-step rule.

The reference's Sedov example (examples/Euler3D/sedov_blast_wave/
sedov_blast_wave.cpp:8-39) deposits the blast energy differently; this
state is the build's own: rho = 1, v = 0 and
p = p_bg + (gamma - 1) E0 exp(-r^2 / (2 sigma^2)) / ((2 pi)^(3/2) sigma^3)
on [-1, 1]^3, with E0 = 1, sigma = 0.0625 and p_bg = 1e-5.
"""
from __future__ import annotations

import numpy as np


def sedov_state(coords: np.ndarray, E0: float = 1.0, sigma: float = 0.0625,
                p_bg: float = 1e-5, gamma: float = 1.4) -> np.ndarray:
    """Conservative state in the reference layout u[e][dof][var] at node
    coordinates [e][dof][3]."""
    r2 = np.sum(coords * coords, axis=-1)
    norm = (2.0 * np.pi) ** 1.5 * sigma ** 3
    p = p_bg + (gamma - 1.0) * E0 * np.exp(-r2 / (2.0 * sigma * sigma)) / norm
    u = np.zeros(coords.shape[:-1] + (5,))
    u[..., 0] = 1.0
    u[..., 4] = p / (gamma - 1.0)
    return u


def sedov_dt(u: np.ndarray, N: int, dx: float, cfl: float = 0.3, gamma: float = 1.4) -> float:
    """cfl dx / (3 (2N + 1) lambda_max), lambda_max the largest |v| + c of u."""
    rho = u[..., 0]
    v = u[..., 1:4] / rho[..., None]
    p = (gamma - 1.0) * (u[..., 4] - 0.5 * rho * np.sum(v * v, axis=-1))
    lam = float(np.max(np.linalg.norm(v, axis=-1) + np.sqrt(gamma * p / rho)))
    return cfl * dx / (3 * (2 * N + 1) * lam)
