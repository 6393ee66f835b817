"""Pins of the oracle's Navier-Stokes (BR1) functions by what the
mathematics fixes, independent of the restatement's own formulas:

* BR1 gradient of a field that is continuous across element faces: the face
  average q_hat equals the trace, so D_hat q plus the lifted face terms is the
  strong derivative D q (summation by parts); with the affine map that is
  (2/h) D q per direction, for the gradient variables (rho, v, T = p/rho).
* Viscous terms of a manufactured smooth state: (NS RHS - Euler RHS) converges
  spectrally to div F_v with Newtonian stress mu (grad v + grad v^T -
  2/3 div v I) and heat flux mu kappa grad T, kappa = gamma / ((gamma-1) Pr).
  Checked for a shear layer (momentum and the v.tau energy term) and for a
  static temperature field (heat conduction), so a dropped term, a wrong
  sign or a wrong factor fails.
* Conservation: the viscous terms leave the quadrature sums of mass,
  momentum and energy unchanged.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

TWO_PI = 2 * math.pi


def state_from_prims(rho, v, p, gamma=1.4):
    m = rho[..., None] * v
    E = p / (gamma - 1) + 0.5 * rho * np.sum(v * v, axis=-1)
    return np.concatenate([rho[..., None], m, E[..., None]], axis=-1)


@pytest.mark.parametrize("dim,N,cells", [(2, 4, (4, 3)), (3, 3, (3, 2, 3))])
def test_br1_gradient_of_continuous_field_is_strong_derivative(dim, N, cells):
    c = O.Case(dim, N, cells, [0.0] * dim, [TWO_PI] * dim, eq=O.NAVIER_STOKES, mu=0.01,
               prandtl=0.72)
    x = c.coordinates()
    s = [np.sin(x[..., k]) for k in range(dim)]
    rho = 1.0 + 0.2 * s[0] * np.cos(x[..., 1])
    v = np.stack([0.3 * np.cos(x[..., (k + 1) % dim]) + 0.1 * s[k] for k in range(dim)], axis=-1)
    p = 1.0 + 0.1 * np.cos(x[..., 0] + x[..., 1])
    u = state_from_prims(rho, v, p)
    g = c.gradients(u)  # (E, nn, nvars, dim)
    q = np.concatenate([rho[..., None], v, (p / rho)[..., None]], axis=-1)  # (E, nn, nvars)
    D = O.basis(N)["D"]
    n = N + 1
    h = [TWO_PI / cells[k] for k in range(dim)]
    shape = (c.E,) + (n,) * dim + (dim + 2,)
    # node index dof = (k*n + j)*n + i  ->  array axes (k, j, i), i fastest
    qq = q.reshape(shape)
    for d in range(dim):
        axis = dim - d  # x_d varies along array axis (dim - d)
        strong = np.moveaxis(np.tensordot(D, np.moveaxis(qq, axis, 0), axes=(1, 0)), 0, axis)
        strong = strong.reshape(c.E, c.ndofs, dim + 2) * (2.0 / h[d])
        got = g[..., d]
        assert np.max(np.abs(got - strong)) <= 1e-11 * np.max(np.abs(strong)), d


def viscous_part(dim, N, cells, u, mu, pr=0.72):
    ns = O.Case(dim, N, cells, [0.0] * dim, [TWO_PI] * dim, eq=O.NAVIER_STOKES, mu=mu,
                prandtl=pr)
    eu = O.Case(dim, N, cells, [0.0] * dim, [TWO_PI] * dim)
    return ns.rhs(u) - eu.rhs(u), ns


@pytest.mark.parametrize("dim", [2, 3])
def test_viscous_terms_of_a_shear_layer(dim):
    """rho = 1, v = (A sin y, 0[, 0]), p = 1: div F_v = (0, -mu A sin y, 0[, 0],
    mu A^2 (cos^2 y - sin^2 y)) -- momentum diffusion and viscous heating."""
    N, cells, mu, A = 7, (6,) * dim, 0.01, 0.5
    c = O.Case(dim, N, cells, [0.0] * dim, [TWO_PI] * dim)
    x = c.coordinates()
    y = x[..., 1]
    rho = np.ones_like(y)
    v = np.zeros(y.shape + (dim,))
    v[..., 0] = A * np.sin(y)
    u = state_from_prims(rho, v, np.ones_like(y))
    got, _ = viscous_part(dim, N, cells, u, mu)
    want = np.zeros_like(got)
    want[..., 1] = -mu * A * np.sin(y)
    want[..., dim + 1] = mu * A * A * (np.cos(y) ** 2 - np.sin(y) ** 2)
    assert np.max(np.abs(got - want)) <= 1e-4 * mu * A


@pytest.mark.parametrize("dim", [2, 3])
def test_heat_conduction_of_a_static_temperature_field(dim):
    """rho = 1, v = 0, p = T = 1 + e cos x: div F_v = (0, 0.., -mu kappa e cos x)."""
    N, cells, mu, eps, pr, gamma = 7, (6,) * dim, 0.02, 0.1, 0.72, 1.4
    c = O.Case(dim, N, cells, [0.0] * dim, [TWO_PI] * dim)
    x = c.coordinates()[..., 0]
    u = state_from_prims(np.ones_like(x), np.zeros(x.shape + (dim,)), 1.0 + eps * np.cos(x))
    got, _ = viscous_part(dim, N, cells, u, mu, pr)
    kappa = gamma / ((gamma - 1) * pr)
    want = np.zeros_like(got)
    want[..., dim + 1] = -mu * kappa * eps * np.cos(x)
    assert np.max(np.abs(got - want)) <= 1e-4 * mu * kappa * eps


def test_viscous_terms_are_conservative():
    dim, N, cells = 3, 3, (3, 3, 2)
    c = O.Case(dim, N, cells, [0.0] * dim, [TWO_PI] * dim)
    rng = np.random.default_rng(4)
    x = c.coordinates()
    rho = 1.0 + 0.1 * np.sin(x[..., 0] + 2 * x[..., 2])
    v = 0.2 * np.stack([np.cos(x[..., 1]), np.sin(x[..., 2]), np.cos(x[..., 0])], axis=-1)
    p = 1.0 + 0.1 * np.cos(x[..., 1] - x[..., 0]) + 0.01 * rng.standard_normal(rho.shape)
    u = state_from_prims(rho, v, p)
    visc, ns = viscous_part(dim, N, cells, u, 0.05)
    w = O.basis(N)["weights"]
    wq = np.einsum("k,j,i->kji", w, w, w).ravel()
    vol = wq[None, :] / np.abs(ns.inverse_jacobian())
    totals = np.einsum("ed,edv->v", vol, visc)
    scale = np.einsum("ed,edv->v", vol, np.abs(visc)).max()
    assert np.max(np.abs(totals)) <= 1e-13 * scale


@pytest.mark.parametrize("dim,N,cells", [(2, 3, (4, 4)), (3, 2, (3, 3, 3))])
def test_viscous_interface_flux_is_the_upper_elements(dim, N, cells):
    """The reference's viscous interface flux (parabolic_viscous.hpp:319-333,
    :571-585) is one-sided: the upper element's (-F_v).n at the face, written
    to both sides.  A change of one element's interior nodes leaves its face
    traces, and with them every neighbour's BR1 gradient and Euler face flux,
    unchanged; it moves the element's own gradient and viscous flux at its
    faces.  So the elements below it (for which it is the upper element) see
    the change, the elements above it do not, and nothing else does.  A
    two-sided average would move both neighbours; the lower element's value
    would move only the ones above."""
    c = O.Case(dim, N, cells, [0.0] * dim, [TWO_PI] * dim, eq=O.NAVIER_STOKES, mu=0.05,
               prandtl=0.72)
    x = c.coordinates()
    rho = 1.0 + 0.1 * np.sin(x[..., 0]) * np.cos(x[..., 1])
    v = np.stack([0.2 * np.cos(x[..., (k + 1) % dim]) for k in range(dim)], axis=-1)
    p = 1.0 + 0.05 * np.sin(x[..., 1])
    u = state_from_prims(rho, v, p)
    n = N + 1
    idx = np.indices((n,) * dim).reshape(dim, -1)[::-1]  # (i, j[, k]) of dof, i fastest
    interior = np.all((idx > 0) & (idx < N), axis=0)
    grid = np.array(cells)
    target = grid // 2
    flat = lambda ijk: int(ijk[0] * cells[1] + ijk[1]) if dim == 2 else \
        int((ijk[0] * cells[1] + ijk[1]) * cells[2] + ijk[2])
    a = flat(target)
    up = u.copy()
    rng = np.random.default_rng(7)
    up[a, interior, 1:dim + 1] += 1e-3 * rng.standard_normal((interior.sum(), dim))
    delta = np.abs(c.rhs(up) - c.rhs(u)).reshape(c.E, -1).max(axis=1)
    below = {flat(target - np.eye(dim, dtype=int)[d]) for d in range(dim)}
    above = {flat(target + np.eye(dim, dtype=int)[d]) for d in range(dim)}
    assert delta[a] > 1e-6
    for e in below:
        assert delta[e] > 1e-7, ("lower neighbour unchanged", e)
    for e in set(range(c.E)) - below - {a}:
        assert delta[e] == 0.0, ("element changed", e, e in above)
