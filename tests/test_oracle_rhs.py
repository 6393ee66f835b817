"""Pins the restatement's right-hand side, indicator, geometry and SSPRK3 by
properties the mathematics fixes: free-stream preservation, conservation,
the discrete entropy identity of the split-form volume integral, entropy
dissipation of the LLF surface flux, closed-form metrics, the indicator's
response to single Legendre modes (including the reference's 3D energy sum,
indicator.hpp:530-544) and third-order convergence of SSPRK3."""
import math

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

from oracle import oracle as O

G = 1.4
TWO_PI = 2 * math.pi


def node_weights(c):
    b = O.basis(c.N)
    w = b["weights"]
    n = c.N + 1
    if c.dim == 2:
        return np.einsum("j,i->ji", w, w).reshape(n * n)
    return np.einsum("k,j,i->kji", w, w, w).reshape(n ** 3)


def entropy_vars(u, dim, g=G):
    rho = u[..., 0]
    m = u[..., 1:dim + 1]
    v2 = np.sum(m * m, axis=-1) / rho ** 2
    p = (g - 1) * (u[..., -1] - 0.5 * rho * v2)
    s = np.log(p) - g * np.log(rho)
    out = np.empty_like(u)
    out[..., 0] = (g - s) / (g - 1) - 0.5 * rho * v2 / p
    out[..., 1:dim + 1] = m / p[..., None]
    out[..., -1] = -rho / p
    return out


def exact_metrics(c, left, right):
    h = [(right[k] - left[k]) / c.cells[k] / 2.0 for k in range(c.dim)]
    Ja = np.zeros((c.E, c.ndofs, c.dim, c.dim))
    for r in range(c.dim):
        Ja[..., r, r] = np.prod([h[k] for k in range(c.dim) if k != r])
    return Ja, np.full((c.E, c.ndofs), 1.0 / np.prod(h)), h


def test_geometry_matches_affine_map():
    left, right = [0.0, -1.0, 0.5], [TWO_PI, 2.0, 1.5]
    c = O.Case(3, 3, (3, 4, 2), left, right)
    Ja_ex, invJ_ex, h = exact_metrics(c, left, right)
    b = O.basis(3)
    x = c.coordinates()
    # element (ie, je, ke) = (1, 2, 1), node (i, j, k) = (2, 0, 3)
    e = (1 * 4 + 2) * 2 + 1
    dof = (3 * 4 + 0) * 4 + 2
    expect = [left[0] + (1 + (1 + b["nodes"][2]) / 2) * 2 * h[0],
              left[1] + (2 + (1 + b["nodes"][0]) / 2) * 2 * h[1],
              left[2] + (1 + (1 + b["nodes"][3]) / 2) * 2 * h[2]]
    np.testing.assert_allclose(x[e, dof], expect, rtol=1e-14, atol=1e-14)
    J = c.jacobian()
    for r in range(3):
        np.testing.assert_allclose(J[..., r, r], h[r], rtol=1e-13)
    np.testing.assert_allclose(c.contravariant(), Ja_ex, rtol=0, atol=1e-11 * np.max(Ja_ex))
    np.testing.assert_allclose(c.inverse_jacobian(), invJ_ex, rtol=1e-12)
    nb = c.left_neighbors()
    assert tuple(nb[0, 0]) == (2, 0, 0) and tuple(nb[0, 1]) == (0, 3, 0) and tuple(nb[0, 2]) == (0, 0, 1)


@pytest.mark.parametrize("dim,N", [(2, 3), (2, 5), (3, 3), (3, 4)])
def test_free_stream_preservation(dim, N):
    cells = (5, 4, 3)[:dim]
    c = O.Case(dim, N, cells, [0.0] * dim, [TWO_PI] * dim)
    u = c.zeros()
    u[..., 0] = 1.3
    u[..., 1] = 0.4
    u[..., 2] = -0.2
    u[..., -1] = 3.1 if dim == 2 else 3.5
    if dim == 3:
        u[..., 3] = 0.25
    du = c.rhs(u)
    assert np.max(np.abs(du)) < 1e-10


@pytest.mark.parametrize("volume", [O.SPLIT_FORM, O.SHOCK_CAPTURING])
def test_conservation_2d(volume):
    c = O.Case(2, 3, (8, 8), [-1, -1], [1, 1], volume=volume, alpha_max=0.5, alpha_min=0.001)
    u = c.init_khi()
    du, alpha = c.rhs(u, with_alpha=True)
    if volume == O.SHOCK_CAPTURING:
        assert np.any(alpha > 0)
    W = node_weights(c)[None, :] / np.abs(c.inverse_jacobian())
    total = np.einsum("en,env->v", W, du)
    scale = np.einsum("en,env->v", W, np.abs(du))
    assert np.all(np.abs(total) <= 1e-13 * scale)


def test_conservation_3d_reference_metrics():
    c = O.Case(3, 3, (4, 3, 3), [0] * 3, [TWO_PI] * 3)
    u = c.init_tgv()
    du = c.rhs(u)
    W = node_weights(c)[None, :] / np.abs(c.inverse_jacobian())
    total = np.einsum("en,env->v", W, du)
    scale = np.einsum("en,env->v", W, np.abs(du))
    assert np.all(np.abs(total) <= 1e-12 * scale)


@pytest.mark.parametrize("dim", [2, 3])
def test_split_form_entropy_identity(dim):
    """Per element, sum_nodes W w(u).vol = sum_dirs sum_lines W_perp (psi_0 - psi_N).Ja_d,
    psi = rho v (Chandrashekar's flux with SBP D_split); exact metrics."""
    N = 3
    n = N + 1
    left, right = [0.0] * dim, [TWO_PI] * dim
    cells = (3, 2, 2)[:dim]
    c = O.Case(dim, N, cells, left, right)
    Ja, invJ, _ = exact_metrics(c, left, right)
    c.set_metrics(Ja, invJ)
    rng = np.random.default_rng(3 + dim)
    u = c.zeros()
    u[..., 0] = 1 + 0.3 * rng.random(u.shape[:2])
    u[..., 1:dim + 1] = 0.2 * rng.standard_normal(u.shape[:2] + (dim,))
    u[..., -1] = 2.5 + 0.1 * rng.random(u.shape[:2])
    vol = c.volume_integral(u)
    w = O.basis(N)["weights"]
    wv = entropy_vars(u, dim)
    W = node_weights(c)
    lhs = np.einsum("n,env->e", W, wv * vol)
    shape = (c.E,) + (n,) * dim + (dim + 2,)
    U = u.reshape(shape)  # axes (e, [k,] j, i, var)
    rhs = np.zeros(c.E)
    for d in range(dim):
        ax = dim - d  # node axis of direction d
        psi = U[..., 1:dim + 1] @ Ja[0, 0, d]  # (rho v).Ja_d at every node
        first = np.take(psi, 0, axis=ax)
        last = np.take(psi, n - 1, axis=ax)
        perp = [w] * (dim - 1)
        Wp = perp[0] if dim == 2 else np.einsum("a,b->ab", w, w)
        rhs += np.sum((first - last).reshape(c.E, -1) * Wp.reshape(-1), axis=1)
    scale = np.einsum("n,env->e", W, np.abs(wv * vol))
    assert np.all(np.abs(lhs - rhs) <= 1e-12 * scale)
    assert np.max(np.abs(rhs)) > 1e-3 * np.max(scale)  # non-degenerate case


def entropy_production(c, u):
    du = c.rhs(u)
    W = node_weights(c)[None, :] / np.abs(c.inverse_jacobian())
    prod = np.einsum("en,env->", W, entropy_vars(u, c.dim) * du)
    scale = np.einsum("en,env->", W, np.abs(entropy_vars(u, c.dim) * du))
    return prod, scale


def test_entropy_conserved_for_continuous_data():
    """EC volume flux + LLF on data continuous across faces: no production."""
    c = O.Case(2, 3, (8, 8), [-1, -1], [1, 1])
    prod, scale = entropy_production(c, c.init_khi())
    assert abs(prod) <= 1e-12 * scale


@pytest.mark.parametrize("dim", [2, 3])
def test_llf_dissipates_entropy_at_jumps(dim):
    """dS/dt = int w . du <= 0 (S = -rho s/(gamma-1)) once faces carry jumps."""
    c = O.Case(dim, 3, (4, 3, 3)[:dim], [0] * dim, [1] * dim)
    rng = np.random.default_rng(dim)
    u = c.zeros()
    u[..., 0] = 1 + 0.3 * rng.random(u.shape[:2])
    u[..., 1:dim + 1] = 0.2 * rng.standard_normal(u.shape[:2] + (dim,))
    u[..., -1] = 2.5 + 0.1 * rng.random(u.shape[:2])
    prod, scale = entropy_production(c, u)
    assert prod < -1e-6 * scale


def single_mode_state(c, axis, amp=0.5):
    """rho = 1, v = 0, p = 1 + amp * P_N(xi_axis) on every element."""
    n = c.N + 1
    x = O.basis(c.N)["nodes"]
    PN = npleg.legval(x, np.eye(n)[c.N])
    u = c.zeros()
    idx = np.arange(c.ndofs)
    node = [(idx // n ** k) % n for k in range(c.dim)]
    p = 1.0 + amp * PN[node[axis]]
    u[..., 0] = 1.0
    u[..., -1] = p[None, :] / (G - 1)
    return u


def test_indicator_constant_and_single_modes():
    c2 = O.Case(2, 3, (2, 2), [-1, -1], [1, 1], volume=O.SHOCK_CAPTURING, alpha_max=0.5)
    u = c2.zeros()
    u[..., 0], u[..., -1] = 1.0, 2.5
    assert np.all(c2.alpha(u) == 0.0)
    for axis in range(2):
        assert np.all(c2.alpha(single_mode_state(c2, axis)) == 0.5)
    c3 = O.Case(3, 3, (2, 2, 2), [-1] * 3, [1] * 3, volume=O.SHOCK_CAPTURING, alpha_max=0.5)
    u = c3.zeros()
    u[..., 0], u[..., -1] = 1.0, 2.5
    assert np.all(c3.alpha(u) == 0.0)
    # the 3D total energy omits the modes (i = N, j < N, k < N): a pure
    # highest mode along xi is invisible, along eta or zeta it is not
    assert np.all(c3.alpha(single_mode_state(c3, 0)) == 0.0)
    assert np.all(c3.alpha(single_mode_state(c3, 1)) == 0.5)
    assert np.all(c3.alpha(single_mode_state(c3, 2)) == 0.5)


def test_indicator_smoothness_ordering():
    """Smooth data below the threshold gives alpha = 0, a jump gives alpha > 0,
    and alpha_max caps the value."""
    c = O.Case(2, 3, (16, 16), [-1, -1], [1, 1], volume=O.SHOCK_CAPTURING, alpha_max=0.002)
    u = c.init_khi()
    a = c.alpha(u)
    assert np.all((a == 0.0) | ((a > 0.0) & (a <= 0.002)))
    assert np.any(a > 0.0) and np.any(a == 0.0)


def test_ssprk3_third_order():
    c = O.Case(3, 3, (2, 2, 2), [0] * 3, [TWO_PI] * 3)
    u0 = c.init_tgv()
    T = 0.06
    ref = c.ssprk3(u0, T / 96, 96)
    errs = []
    for nsteps in (6, 12):
        errs.append(np.linalg.norm(c.ssprk3(u0, T / nsteps, nsteps) - ref))
    ratio = errs[0] / errs[1]
    assert 6.5 < ratio < 9.5, ratio


def test_ssprk3_constant_state_is_steady():
    c = O.Case(2, 2, (3, 3), [-1, -1], [1, 1])
    u = c.zeros()
    u[..., 0], u[..., 1], u[..., 2], u[..., 3] = 1.0, 0.1, -0.2, 2.0
    out = c.ssprk3(u, 1e-2, 5)
    assert np.max(np.abs(out - u)) < 1e-12
