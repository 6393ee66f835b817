"""Pins the restatement's CentralFlux (numerical_flux.hpp:230-258 with
numerical_flux_detail.hpp:23-32) and its wiring into the split-form volume
integral and the interface flux.

* Pointwise: 0.5 (F(u_L).n + F(u_R).n) against the textbook Euler flux,
  consistency and symmetry.
* RHS: with the central two-point flux as the volume flux, the split form
  sum_m D_split(i, m) 0.5 (Ft_i + Ft_m) equals (D Ft)_i - (B_ii / w_i) Ft_i
  for the contravariant flux Ft = Ja.F (the rows of D sum to zero and
  D_split = 2 D - W^-1 B); with the central surface flux the surface term
  adds (B_ii / w_i) f*.  So the RHS is the strong-form DGSEM
  -(1/J) [(D Ft)_i + (B_ii / w_i)(f* - Ft_i)], which the test writes out in
  numpy from the textbook flux on an affine periodic 2D mesh.  A volume
  flux left at Chandrashekar, a missing factor 1/2 or a surface flux left
  at LLF fails it.
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_flux import G, cons, euler_flux_n, random_state


def test_central_flux_pointwise():
    rng = np.random.default_rng(5)
    for _ in range(200):
        ul, ur = random_state(rng, 2), random_state(rng, 2)
        n = rng.standard_normal(2)
        f = O.flux_central(2, G, ul, ur, n)
        ref = 0.5 * (euler_flux_n(2, ul, n) + euler_flux_n(2, ur, n))
        assert np.allclose(f, ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())
        # symmetric in its states, consistent with the physical flux
        assert np.array_equal(f, O.flux_central(2, G, ur, ul, n))
        fu = O.flux_central(2, G, ul, ul, n)
        assert np.allclose(fu, euler_flux_n(2, ul, n), rtol=1e-14, atol=1e-15)
        # at one state LLF has no dissipation and is the central flux
        assert np.allclose(O.flux_llf(2, G, ul, ul, n), fu, rtol=1e-14, atol=1e-15)


def strong_form_rhs(u, N, cells, left, right, g=G):
    """-(1/J)[(D Ft)_i + (B_ii / w_i)(f* - Ft_i)] on an affine periodic 2D
    mesh, from the textbook flux; u in the reference layout
    [e][dof][var], e = ie * ny + je, dof = j * n + i."""
    b = O.basis(N)
    D, w = b["D"], b["weights"]
    n = N + 1
    nx, ny = cells
    hx, hy = (right[0] - left[0]) / nx, (right[1] - left[1]) / ny
    ja = (hy / 2, hx / 2)                  # Ja^1 = (hy/2, 0), Ja^2 = (0, hx/2)
    invJ = 4.0 / (hx * hy)
    U = u.reshape(nx, ny, n, n, 4)         # [ie][je][j][i][var]
    Fx = np.zeros_like(U)
    Fy = np.zeros_like(U)
    for idx in np.ndindex(nx, ny, n, n):
        Fx[idx] = ja[0] * euler_flux_n(2, U[idx], np.array([1.0, 0.0]), g)
        Fy[idx] = ja[1] * euler_flux_n(2, U[idx], np.array([0.0, 1.0]), g)
    vol = np.einsum("im,xyjmv->xyjiv", D, Fx) + np.einsum("jm,xymiv->xyjiv", D, Fy)
    surf = np.zeros_like(U)
    # x faces: the right neighbour's first column, the left neighbour's last
    Fx_r = np.roll(Fx, -1, axis=0)[:, :, :, 0, :]
    Fx_l = np.roll(Fx, 1, axis=0)[:, :, :, n - 1, :]
    surf[:, :, :, n - 1, :] += (0.5 * (Fx[:, :, :, n - 1, :] + Fx_r) - Fx[:, :, :, n - 1, :]) / w[n - 1]
    surf[:, :, :, 0, :] -= (0.5 * (Fx_l + Fx[:, :, :, 0, :]) - Fx[:, :, :, 0, :]) / w[0]
    Fy_u = np.roll(Fy, -1, axis=1)[:, :, 0, :, :]
    Fy_d = np.roll(Fy, 1, axis=1)[:, :, n - 1, :, :]
    surf[:, :, n - 1, :, :] += (0.5 * (Fy[:, :, n - 1, :, :] + Fy_u) - Fy[:, :, n - 1, :, :]) / w[n - 1]
    surf[:, :, 0, :, :] -= (0.5 * (Fy_d + Fy[:, :, 0, :, :]) - Fy[:, :, 0, :, :]) / w[0]
    return (-invJ * (vol + surf)).reshape(u.shape)


def affine_case(N, cells, left, right, **kw):
    c = O.Case(2, N, cells, left, right, **kw)
    h = [(right[k] - left[k]) / cells[k] / 2.0 for k in range(2)]
    Ja = np.zeros((c.E, c.ndofs, 2, 2))
    Ja[..., 0, 0], Ja[..., 1, 1] = h[1], h[0]
    c.set_metrics(Ja, np.full((c.E, c.ndofs), 1.0 / (h[0] * h[1])))
    return c


def smooth_state(x):
    X, Y = x[..., 0], x[..., 1]
    rho = 1.0 + 0.3 * np.sin(np.pi * X + 0.2) * np.cos(np.pi * Y)
    vx = 0.4 + 0.2 * np.cos(np.pi * Y + 0.5)
    vy = -0.3 + 0.1 * np.sin(np.pi * X)
    p = 1.0 + 0.2 * np.cos(np.pi * (X + Y))
    return np.stack([rho, rho * vx, rho * vy, p / (G - 1) + 0.5 * rho * (vx * vx + vy * vy)], -1)


@pytest.mark.parametrize("N,cells", [(2, (3, 4)), (3, (4, 3)), (5, (2, 3))])
def test_central_split_form_is_the_strong_form(N, cells):
    left, right = (-1.0, -1.0), (1.0, 1.0)
    c = affine_case(N, cells, left, right, volume_flux=O.VOLUME_CENTRAL, surface_flux=O.FLUX_CENTRAL)
    rng = np.random.default_rng(N)
    # a smooth state with element-wise jumps, so that the surface terms count
    u = smooth_state(c.coordinates()) * (1 + 0.05 * rng.random((c.E, 1, 1)))
    du = c.rhs(u)
    ref = strong_form_rhs(u, N, cells, left, right)
    err = np.linalg.norm(du - ref) / np.linalg.norm(ref)
    assert err < 1e-13, err
    # the planted mistakes the identity catches
    chandra = affine_case(N, cells, left, right, surface_flux=O.FLUX_CENTRAL).rhs(u)
    assert np.linalg.norm(chandra - ref) / np.linalg.norm(ref) > 1e-4
    llf = affine_case(N, cells, left, right, volume_flux=O.VOLUME_CENTRAL).rhs(u)
    assert np.linalg.norm(llf - ref) / np.linalg.norm(ref) > 1e-4


def test_central_fluxes_preserve_a_constant_state():
    c = O.Case(2, 3, (3, 3), (0.0, 0.0), (1.0, 1.0), volume_flux=O.VOLUME_CENTRAL,
               surface_flux=O.FLUX_CENTRAL)
    u = np.broadcast_to(cons(2, 1.2, [0.3, -0.2], 0.9), (c.E, c.ndofs, 4)).copy()
    assert np.abs(c.rhs(u)).max() < 1e-12
