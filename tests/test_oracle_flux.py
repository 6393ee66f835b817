"""Pins the restatement's pointwise physics (logarithmic mean, physical flux,
Chandrashekar, LLF, HLLC) by closed forms, consistency, symmetry, Tadmor's
entropy-conservation identity and rotational invariance."""
import math

import numpy as np
import pytest

from oracle import oracle as O

G = 1.4


def cons(dim, rho, v, p, g=G):
    v = np.asarray(v, dtype=float)
    return np.array([rho, *(rho * v), p / (g - 1) + 0.5 * rho * v @ v])


def euler_flux_n(dim, u, n, g=G):
    """Textbook Euler flux F(u).n, written from the equations."""
    rho, m, E = u[0], u[1:dim + 1], u[-1]
    v = m / rho
    p = (g - 1) * (E - 0.5 * rho * v @ v)
    vn = v @ n
    return np.array([rho * vn, *(m * vn + p * np.asarray(n)), (E + p) * vn])


def entropy_vars(dim, u, g=G):
    rho, v = u[0], u[1:dim + 1] / u[0]
    p = (g - 1) * (u[-1] - 0.5 * rho * v @ v)
    s = math.log(p) - g * math.log(rho)
    return np.array([(g - s) / (g - 1) - 0.5 * rho * v @ v / p, *(rho * v / p), -rho / p])


def random_state(rng, dim, scale=1.0):
    return cons(dim, 0.5 + rng.random(), scale * rng.standard_normal(dim), 0.5 + rng.random())


def test_ln_mean_closed_form_and_branches():
    for x, y in [(1.0, 2.0), (0.3, 7.0), (2.5, 2.49), (1e-3, 1e3)]:
        exact = (y - x) / math.log1p((y - x) / x)
        assert abs(O.ln_mean(x, y) - exact) <= 2e-15 * exact
        assert O.ln_mean(x, y) == O.ln_mean(y, x) or abs(O.ln_mean(x, y) - O.ln_mean(y, x)) < 1e-15 * exact
    for x in [1e-7, 0.37, 1.0, 42.0]:
        assert abs(O.ln_mean(x, x) - x) <= 1e-16 * x
    # series branch (f2 < 1e-4) agrees with the closed form near the switch
    for r in [1.0 + 1e-6, 1.01, 1.0199, 1.0201, 1.03]:
        x, y = 1.0, r
        exact = (y - x) / math.log1p(y - x)
        assert abs(O.ln_mean(x, y) - exact) <= 1e-13 * exact


@pytest.mark.parametrize("dim", [2, 3])
def test_physical_flux_matches_textbook(dim):
    rng = np.random.default_rng(1)
    for _ in range(5):
        u, n = random_state(rng, dim), rng.standard_normal(dim)
        np.testing.assert_allclose(O.physical_flux(dim, G, u, n), euler_flux_n(dim, u, n), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_chandrashekar_consistency_symmetry_entropy(dim):
    rng = np.random.default_rng(2 + dim)
    for _ in range(20):
        a, b, n = random_state(rng, dim), random_state(rng, dim), rng.standard_normal(dim)
        np.testing.assert_allclose(O.flux_chandrashekar(dim, G, a, a, n), euler_flux_n(dim, a, n),
                                   rtol=1e-12, atol=1e-12)
        fab = O.flux_chandrashekar(dim, G, a, b, n)
        np.testing.assert_allclose(fab, O.flux_chandrashekar(dim, G, b, a, n), rtol=1e-14, atol=1e-14)
        # Tadmor: (w_b - w_a) . f = (psi_b - psi_a) . n with psi = rho v
        lhs = (entropy_vars(dim, b) - entropy_vars(dim, a)) @ fab
        rhs = (b[1:dim + 1] - a[1:dim + 1]) @ n
        assert abs(lhs - rhs) <= 1e-11 * (1 + abs(rhs))


@pytest.mark.parametrize("dim", [2, 3])
def test_llf_consistency_and_dissipation(dim):
    rng = np.random.default_rng(7)
    for _ in range(10):
        a, b, n = random_state(rng, dim), random_state(rng, dim), rng.standard_normal(dim)
        np.testing.assert_allclose(O.flux_llf(dim, G, a, a, n), euler_flux_n(dim, a, n), rtol=1e-13, atol=1e-13)
        lam = 0.0
        va, vb = a[1:dim + 1] / a[0], b[1:dim + 1] / b[0]
        pa = (G - 1) * (a[-1] - 0.5 * a[0] * va @ va)
        pb = (G - 1) * (b[-1] - 0.5 * b[0] * vb @ vb)
        lam = max(abs(va @ n), abs(vb @ n)) + max(math.sqrt(G * pa / a[0]), math.sqrt(G * pb / b[0])) * np.linalg.norm(n)
        central = 0.5 * (euler_flux_n(dim, a, n) + euler_flux_n(dim, b, n))
        np.testing.assert_allclose(O.flux_llf(dim, G, a, b, n), central - 0.5 * lam * (b - a), rtol=1e-12, atol=1e-12)


def rotation(dim, rng):
    q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
    return q


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_consistency_rotation_upwinding(dim):
    rng = np.random.default_rng(11)
    for _ in range(20):
        a, b, n = random_state(rng, dim, 2.0), random_state(rng, dim, 2.0), rng.standard_normal(dim)
        np.testing.assert_allclose(O.flux_hllc(dim, G, a, a, n), euler_flux_n(dim, a, n), rtol=1e-12, atol=1e-12)
        R = rotation(dim, rng)

        def rot(u):
            return np.array([u[0], *(R @ u[1:dim + 1]), u[-1]])

        f = O.flux_hllc(dim, G, a, b, n)
        fr = O.flux_hllc(dim, G, rot(a), rot(b), R @ n)
        np.testing.assert_allclose(fr, rot(f), rtol=1e-11, atol=1e-11)
        # scaling the normal scales the flux
        np.testing.assert_allclose(O.flux_hllc(dim, G, a, b, 3.0 * n), 3.0 * f, rtol=1e-12, atol=1e-12)
    # supersonic flow along n: pure upwinding
    n = np.eye(dim)[0]
    a = cons(dim, 1.0, 10.0 * n, 1.0)
    b = cons(dim, 0.5, 12.0 * n, 0.7)
    np.testing.assert_allclose(O.flux_hllc(dim, G, a, b, n), euler_flux_n(dim, a, n), rtol=1e-13)
    # with the normal reversed the same flow is upwinded from the other side
    np.testing.assert_allclose(O.flux_hllc(dim, G, b, a, -n), euler_flux_n(dim, a, -n), rtol=1e-13)


def normal_shock(dim, rng, mach, sign):
    """An isolated shock of the family moving along +n (sign 1) or -n (sign -1)
    relative to the fluid ahead of it, from the normal-shock relations, with
    a tangential velocity that the shock does not change.  Returns (u_l, u_r,
    n, shock speed along n)."""
    n = rng.standard_normal(dim)
    n /= np.linalg.norm(n)
    t = rng.standard_normal(dim)
    t -= (t @ n) * n
    rho_a, p_a, vn_a = 0.8 + 0.4 * rng.random(), 0.7 + 0.6 * rng.random(), rng.standard_normal()
    c_a = math.sqrt(G * p_a / rho_a)
    ratio = (G + 1) * mach ** 2 / ((G - 1) * mach ** 2 + 2)
    rho_b = rho_a * ratio
    p_b = p_a * (1 + 2 * G / (G + 1) * (mach ** 2 - 1))
    S = vn_a + sign * mach * c_a
    vn_b = S - (S - vn_a) / ratio
    ahead = cons(dim, rho_a, vn_a * n + t, p_a)
    behind = cons(dim, rho_b, vn_b * n + t, p_b)
    # the shock moves into the state ahead of it
    return (behind, ahead, n, S) if sign > 0 else (ahead, behind, n, S)


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_resolves_isolated_shocks_and_contacts(dim):
    """With the reference's wave-speed bounds (min/max of the one-sided speeds
    and the Roe-averaged ones, numerical_flux.hpp:331-335) the Roe speed of an
    isolated shock is its Rankine-Hugoniot speed, and the HLLC middle state
    then collapses onto the exact Riemann solution: the flux through the
    interface is the physical flux of the state on the shock's side of
    x/t = 0.  Contacts (equal pressure and normal velocity) are resolved
    exactly too.  Plain one-sided bounds (v -/+ c) or a wrong star state break
    the shock case."""
    rng = np.random.default_rng(5)
    for _ in range(12):
        for mach in (1.3, 2.5, 7.0):
            for sign in (1, -1):
                ul, ur, n, S = normal_shock(dim, rng, mach, sign)
                upwind = ul if S > 0 else ur
                f = O.flux_hllc(dim, G, ul, ur, n)
                np.testing.assert_allclose(f, euler_flux_n(dim, upwind, n), rtol=1e-11, atol=1e-11)
                # the interface normal's length scales the flux
                np.testing.assert_allclose(O.flux_hllc(dim, G, ul, ur, 2.5 * n), 2.5 * f,
                                           rtol=1e-12, atol=1e-12)
        n = rng.standard_normal(dim)
        n /= np.linalg.norm(n)
        vn, p = rng.standard_normal(), 0.5 + rng.random()
        tl, tr = rng.standard_normal(dim), rng.standard_normal(dim)
        tl -= (tl @ n) * n
        tr -= (tr @ n) * n
        ul = cons(dim, 0.5 + rng.random(), vn * n + tl, p)
        ur = cons(dim, 1.5 + rng.random(), vn * n + tr, p)
        upwind = ul if vn > 0 else ur
        np.testing.assert_allclose(O.flux_hllc(dim, G, ul, ur, n), euler_flux_n(dim, upwind, n),
                                   rtol=1e-12, atol=1e-12)
