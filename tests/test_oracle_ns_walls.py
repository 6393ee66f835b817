"""Pins of the oracle's Navier-Stokes adiabatic walls (no slip and slip,
2D), by what the physics and the reference's definitions fix:

* a fluid at rest with uniform density and temperature is steady between
  any mix of the walls (the wall keeps the inner density and temperature in
  its gradient value; dropping either would create a spurious gradient);
* uniform flow along slip walls is steady, and is not along no-slip walls;
* adiabatic walls do no work and pass no heat: in a closed box of walls the
  mass and total-energy integrals have zero rate for any state;
* the no-slip wall shear: for rho = p = 1 and v = (0, cos(pi x / 2)) between
  no-slip walls at x = -1, 1 (y periodic), the inviscid terms vanish and the
  y-momentum integral changes at the rate mu [dv/dx] = -mu pi per unit
  length in y (spectral convergence of the BR1 gradient trace); slip walls
  carry no shear (zero rate);
* the slip wall removes only the normal velocity from the gradient value: for
  v = (U, 0) its gradients equal the no-slip wall's, and for v = (0, V) the
  velocity gradient is zero.
"""
import numpy as np
import pytest

from oracle import oracle as O

GAMMA, MU, PR = 1.4, 5.0e-3, 0.73
NOSLIP, SLIP = O.BC_NS_NOSLIP_WALL, O.BC_NS_SLIP_WALL


def ns_box(N, cells, faces, left=(-1.0, -0.5), right=(1.0, 0.5), mu=MU):
    periodic = [int(faces[0] == O.BC_PERIODIC), int(faces[2] == O.BC_PERIODIC)]
    c = O.Case(2, N, cells, list(left), list(right), periodic=periodic, eq=O.NAVIER_STOKES,
               gamma=GAMMA, mu=mu, prandtl=PR)
    for f, k in enumerate(faces):
        if k != O.BC_PERIODIC:
            c.set_boundary(f, k)
    return c


def state(c, rho, vx, vy, p):
    u = np.zeros((c.E, c.ndofs, 4))
    u[..., 0] = rho
    u[..., 1] = rho * vx
    u[..., 2] = rho * vy
    u[..., 3] = p / (GAMMA - 1) + 0.5 * rho * (vx * vx + vy * vy)
    return u


def volumes(c, N):
    w = O.basis(N)["weights"]
    return np.einsum("j,i->ji", w, w).ravel()[None, :] / np.abs(c.inverse_jacobian())


@pytest.mark.parametrize("faces", [[NOSLIP] * 4, [SLIP] * 4, [NOSLIP, NOSLIP, NOSLIP, SLIP],
                                   [SLIP, NOSLIP, O.BC_PERIODIC, O.BC_PERIODIC]])
def test_rest_state_is_steady(faces):
    c = ns_box(3, (5, 4), faces)
    u = state(c, 1.3, 0.0, 0.0, 0.8)
    assert np.max(np.abs(c.rhs(u))) <= 1e-12


def test_tangential_flow_steady_only_along_slip_walls():
    faces_slip = [SLIP, SLIP, O.BC_PERIODIC, O.BC_PERIODIC]
    faces_noslip = [NOSLIP, NOSLIP, O.BC_PERIODIC, O.BC_PERIODIC]
    for faces, steady in ((faces_slip, True), (faces_noslip, False)):
        c = ns_box(3, (5, 4), faces)
        du = c.rhs(state(c, 1.1, 0.0, 0.4, 1.0))
        assert (np.max(np.abs(du)) <= 1e-12) == steady, faces


@pytest.mark.parametrize("faces", [[NOSLIP] * 4, [SLIP] * 4, [NOSLIP, NOSLIP, NOSLIP, SLIP]])
def test_closed_box_conserves_mass_and_energy(faces):
    N = 3
    c = ns_box(N, (6, 5), faces)
    x = c.coordinates()
    rng = np.random.default_rng(5)
    rho = 1.0 + 0.2 * np.sin(2 * x[..., 0] + 1) * np.cos(3 * x[..., 1]) + 0.01 * rng.random(x.shape[:-1])
    u = state(c, rho, 0.3 * np.cos(x[..., 1] * 3), 0.2 * np.sin(2 * x[..., 0]),
              1.0 + 0.1 * np.cos(x[..., 0] - x[..., 1]))
    du = c.rhs(u)
    vol = volumes(c, N)
    for var in (0, 3):
        rate = np.sum(vol * du[..., var])
        assert abs(rate) <= 1e-13 * np.sum(vol * np.abs(du[..., var])), var


@pytest.mark.parametrize("wall,expected", [(NOSLIP, -MU * np.pi), (SLIP, 0.0)])
def test_wall_shear_closed_form(wall, expected):
    N, cells = 6, (6, 2)
    c = ns_box(N, cells, [wall, wall, O.BC_PERIODIC, O.BC_PERIODIC])
    x = c.coordinates()
    u = state(c, 1.0, 0.0, np.cos(np.pi * x[..., 0] / 2), 1.0)
    du = c.rhs(u)
    vol = volumes(c, N)
    rate = np.sum(vol * du[..., 2])  # per unit length in y: the box is 1 long
    if expected:
        assert abs(rate - expected) <= 1e-6 * abs(expected), rate
    else:
        # slip walls: only the interior viscous terms act, which conserve
        assert abs(rate) <= 1e-13 * np.sum(vol * np.abs(du[..., 2])), rate
    assert abs(np.sum(vol * du[..., 3])) <= 1e-12 * np.sum(vol * np.abs(du[..., 3]))


def test_slip_wall_gradient_drops_only_the_normal_velocity():
    faces_slip = [SLIP, SLIP, O.BC_PERIODIC, O.BC_PERIODIC]
    faces_noslip = [NOSLIP, NOSLIP, O.BC_PERIODIC, O.BC_PERIODIC]
    cs, cn = ns_box(3, (4, 3), faces_slip), ns_box(3, (4, 3), faces_noslip)
    x = cs.coordinates()
    p = 1.0 + 0.1 * np.cos(x[..., 0] + 2 * x[..., 1])
    u = state(cs, 1.2, 0.5, 0.0, p)
    gs, gn = cs.gradients(u), cn.gradients(u)
    assert np.max(np.abs(gs - gn)) <= 1e-13 * np.max(np.abs(gn))
    g = cs.gradients(state(cs, 1.2, 0.0, 0.5, p))
    assert np.max(np.abs(g[..., 1:3, :])) <= 1e-12


# ---- Dirichlet faces under Navier-Stokes (dirichlet_boundary.hpp:122-194) ----

def face_traces(c, u, face, N, cells):
    """Inner traces of a face in boundary-node order: (element position along
    the face) * (N+1) + node position."""
    n, (nx, ny) = N + 1, cells
    out = []
    count = ny if face < 2 else nx
    for idx in range(count):
        for q in range(n):
            if face < 2:
                e, dof = (0 if face == 0 else nx - 1) * ny + idx, q * n + (0 if face == 0 else n - 1)
            else:
                e, dof = idx * ny + (0 if face == 2 else ny - 1), (0 if face == 2 else n - 1) * n + q
            out.append(u[e, dof])
    return np.array(out)


def test_dirichlet_free_stream_is_steady_under_ns():
    """Dirichlet faces carry a gradient value (the outer state's variables):
    a uniform flow with the same Dirichlet state everywhere is steady."""
    c = ns_box(3, (5, 4), [O.BC_DIRICHLET] * 4)
    u = state(c, 1.2, 0.3, -0.2, 0.9)
    for f in range(4):
        c.set_boundary(f, O.BC_DIRICHLET, state=u[0, 0])
    assert np.max(np.abs(c.rhs(u))) <= 1e-12


def test_dirichlet_steady_heat_conduction():
    """Gas at rest with uniform pressure and a linear temperature between two
    Dirichlet faces holding the exact states: the heat flux is constant, so
    the RHS vanishes -- which needs the Dirichlet viscous face value to carry
    the energy flux (the adiabatic wall zeroes it)."""
    N, cells = 4, (4, 3)
    c = ns_box(N, cells, [O.BC_DIRICHLET, O.BC_DIRICHLET, O.BC_PERIODIC, O.BC_PERIODIC], mu=0.05)
    x = c.coordinates()
    T = 1.0 + 0.2 * x[..., 0]
    u = state(c, 1.0 / T, 0.0, 0.0, 1.0)
    for f in (0, 1):
        c.set_boundary(f, O.BC_DIRICHLET, values=face_traces(c, u, f, N, cells))
    du = c.rhs(u)
    scale = 0.05 * 1.4 / 0.4 / PR * 0.2  # kappa dT/dx
    assert np.max(np.abs(du)) <= 1e-11 * scale / 0.05
    # the adiabatic wall in its place does change the energy
    c2 = ns_box(N, cells, [NOSLIP, O.BC_DIRICHLET, O.BC_PERIODIC, O.BC_PERIODIC], mu=0.05)
    c2.set_boundary(1, O.BC_DIRICHLET, values=face_traces(c2, u, 1, N, cells))
    assert np.max(np.abs(c2.rhs(u)[..., 3])) > 1e-3 * scale


def test_dirichlet_gradient_value_is_the_outer_state():
    """Changing a Dirichlet state changes the BR1 gradient only at that face's
    nodes, by the lift of the difference of the outer gradient variables:
    the same factor multiplies every variable's difference."""
    N, cells = 3, (4, 3)
    faces = [O.BC_DIRICHLET, O.BC_DIRICHLET, O.BC_PERIODIC, O.BC_PERIODIC]
    c = ns_box(N, cells, faces)
    x = c.coordinates()
    u = state(c, 1.0 + 0.1 * np.sin(x[..., 0]), 0.2, 0.1 * np.cos(x[..., 1]), 1.0)
    s1 = np.array([1.1, 0.2, 0.0, 2.5])
    s2 = np.array([0.9, -0.1, 0.3, 2.2])
    c.set_boundary(1, O.BC_DIRICHLET, state=s1)
    c.set_boundary(0, O.BC_DIRICHLET, state=s1)
    g1 = c.gradients(u)
    c.set_boundary(0, O.BC_DIRICHLET, state=s2)
    g2 = c.gradients(u)

    def q(s):
        rho, vx, vy = s[0], s[1] / s[0], s[2] / s[0]
        p = (GAMMA - 1) * (s[3] - 0.5 * rho * (vx * vx + vy * vy))
        return np.array([rho, vx, vy, p / rho])

    dq = q(s2) - q(s1)
    dg = g2 - g1
    n = N + 1
    on_face = np.zeros(dg.shape[:2], dtype=bool)
    for je in range(cells[1]):
        on_face[je, [j * n for j in range(n)]] = True  # ie = 0, node i = 0
    assert np.max(np.abs(dg[~on_face])) <= 1e-13
    ratio = dg[on_face] / dq[None, :, None]  # (nodes, var, dim)
    assert np.max(np.abs(ratio[:, 0, 0])) > 1.0
    assert np.max(np.abs(ratio - ratio[:, :1, :])) <= 1e-12 * np.max(np.abs(ratio))


# ---- MixedDirichletSlipWallBC (dirichlet_boundary.hpp:198-275), Euler ----

def test_mixed_dirichlet_slip_wall_selects_per_node():
    """Where the flag is set the node takes the Dirichlet state, elsewhere the
    slip-wall reflection: the RHS equals the all-Dirichlet RHS at flagged
    boundary nodes and the all-wall RHS everywhere else, bit for bit."""
    N, cells = 3, (5, 4)
    n = N + 1

    def euler(bottom, **kw):
        c = O.Case(2, N, cells, [-1.0, -0.5], [1.0, 0.75], periodic=[1, 0])
        c.set_boundary(3, O.BC_SLIP_WALL)
        c.set_boundary(2, bottom, **kw)
        return c

    cw = euler(O.BC_SLIP_WALL)
    x = cw.coordinates()
    u = state(cw, 1.0 + 0.2 * np.sin(x[..., 0]), 0.3, 0.2 * np.cos(x[..., 0]), 1.0 + 0.1 * x[..., 1])
    vals = np.tile([1.3, 0.4, 0.1, 2.7], (cells[0] * n, 1))
    vals[:, 0] += 0.05 * np.arange(len(vals)) / len(vals)
    flags = (np.arange(len(vals)) % 3 == 0).astype(float)
    cd = euler(O.BC_DIRICHLET, values=vals)
    cm = euler(O.BC_MIXED_DIRICHLET_SLIP_WALL, values=np.column_stack([vals, flags]))
    rw, rd, rm = cw.rhs(u), cd.rhs(u), cm.rhs(u)
    expect = rw.copy()
    for ie in range(cells[0]):
        for q in range(n):
            if flags[ie * n + q]:
                e = ie * cells[1]  # je = 0
                expect[e, q] = rd[e, q]  # bottom row node j = 0 has dof = q
    assert np.array_equal(rm, expect)
    assert not np.array_equal(rm, rw) and not np.array_equal(rm, rd)
