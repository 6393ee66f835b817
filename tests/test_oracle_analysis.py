"""The oracle's analysis reduction (orc_analyze) pinned by closed forms:
total volume of the box, kinetic energy and entropy of states whose
integrals are known exactly (the GLL quadrature integrates the trigonometric
polynomials below exactly to rounding at these resolutions), the non-finite
count, and the largest signal speed."""
import math

import numpy as np

from oracle import oracle as O


def test_volume_kinetic_energy_entropy_closed_forms():
    c = O.Case(3, 4, (4, 4, 4), [0, 0, 0], [2 * math.pi] * 3)
    x = c.coordinates()
    rho = np.ones(x.shape[:-1])
    v = np.stack([np.sin(x[..., 0]), np.zeros_like(rho), np.zeros_like(rho)], axis=-1)
    p = np.full_like(rho, 2.0)
    E = p / 0.4 + 0.5 * rho * np.sum(v * v, axis=-1)
    u = np.concatenate([rho[..., None], rho[..., None] * v, E[..., None]], axis=-1)
    vol, ke, ent, bad, vmax = c.analyze(u)
    L3 = (2 * math.pi) ** 3
    assert abs(vol - L3) <= 1e-12 * L3
    assert abs(ke - 0.25 * L3) <= 1e-10 * L3  # <sin^2>/2 = 1/4
    s = math.log(2.0) - 1.4 * math.log(1.0)
    assert abs(ent - (-s / 0.4) * L3) <= 1e-12 * L3
    assert bad == 0
    assert abs(vmax - (1.0 + math.sqrt(1.4 * 2.0))) <= 0.01  # max |sin| at a node
    assert abs(ke - c.kinetic_energy(u)) <= 1e-14 * ke


def test_non_finite_nodes_are_counted_and_skipped():
    c = O.Case(2, 3, (2, 2), [0, 0], [1, 1])
    u = np.zeros((c.E, c.ndofs, 4))
    u[..., 0] = 1.0
    u[..., 3] = 2.5
    vol0, ke0, ent0, bad0, _ = c.analyze(u)
    u[1, 3, 2] = np.nan
    u[2, 0, 0] = np.inf
    vol, ke, ent, bad, _ = c.analyze(u)
    assert bad == 2 and bad0 == 0
    assert vol < vol0
