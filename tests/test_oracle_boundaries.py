"""Pins of the oracle's 2D boundary conditions (slip wall, outflow,
Dirichlet), by properties the boundary fluxes must have:

* Dirichlet with the free-stream state itself, on every face, keeps a
  uniform flow steady (RHS zero to rounding) -- as periodicity does;
* outflow is the same for a uniform state (its outer state is the line
  mean, equal to the state);
* a slip wall carries no mass: for any state, the mass flux through a wall
  vanishes, so the quadrature total of d rho/dt over the domain is zero
  when every face is a wall (and momentum tangential to a wall is only
  changed by the interior);
* the reflected state has the normal velocity reversed and the same
  density, tangential velocity and energy: a state at rest next to a wall
  gives the pressure flux p n only;
* per-node Dirichlet states equal to a constant reproduce the constant-state
  Dirichlet flux bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle as O

GAMMA = 1.4


def box(N=3, cells=(6, 5), periodic=(0, 0)):
    return O.Case(2, N, cells, [-1.0, -1.0], [1.0, 1.0], periodic=list(periodic))


def uniform(c, rho=1.2, vx=0.3, vy=-0.2, p=0.9):
    u = np.zeros((c.E, c.ndofs, 4))
    u[..., 0] = rho
    u[..., 1] = rho * vx
    u[..., 2] = rho * vy
    u[..., 3] = p / (GAMMA - 1) + 0.5 * rho * (vx * vx + vy * vy)
    return u


def random_state(c, seed=0):
    rng = np.random.default_rng(seed)
    x = c.coordinates()
    rho = 1.0 + 0.2 * np.sin(3 * x[..., 0]) * np.cos(2 * x[..., 1]) + 0.01 * rng.random(x.shape[:-1])
    vx = 0.3 * np.cos(2 * x[..., 1]) + 0.05 * rng.random(rho.shape)
    vy = 0.2 * np.sin(3 * x[..., 0])
    p = 1.0 + 0.1 * np.cos(x[..., 0] + x[..., 1])
    u = np.stack([rho, rho * vx, rho * vy, p / (GAMMA - 1) + 0.5 * rho * (vx * vx + vy * vy)], -1)
    return u


@pytest.mark.parametrize("kind", [O.BC_DIRICHLET, O.BC_OUTFLOW])
def test_uniform_flow_is_steady(kind):
    c = box()
    u = uniform(c)
    for face in range(4):
        c.set_boundary(face, kind, state=u[0, 0])
    du = c.rhs(u)
    assert np.max(np.abs(du)) <= 1e-12


def test_slip_walls_conserve_mass():
    c = box(periodic=(0, 0))
    for face in range(4):
        c.set_boundary(face, O.BC_SLIP_WALL)
    u = random_state(c)
    du = c.rhs(u)
    w = O.basis(3)["weights"]
    wq = np.einsum("j,i->ji", w, w).ravel()
    vol = wq[None, :] / np.abs(c.inverse_jacobian())
    mass_rate = np.sum(vol * du[..., 0])
    assert abs(mass_rate) <= 1e-13 * np.sum(vol * np.abs(du[..., 0]))
    # with x-walls only and y periodic, y-momentum is conserved too
    c2 = box(periodic=(0, 1))
    c2.set_boundary(0, O.BC_SLIP_WALL)
    c2.set_boundary(1, O.BC_SLIP_WALL)
    du2 = c2.rhs(u)
    for var in (0, 2):
        rate = np.sum(vol * du2[..., var])
        assert abs(rate) <= 1e-13 * np.sum(vol * np.abs(du2[..., var])), var


def test_slip_wall_at_rest_gives_pressure_only():
    """Fluid at rest with a pressure field: the wall's reflected state is the
    inner state, so the boundary flux is the consistent flux F(u).n = (0, p n,
    0); the result equals that of a Dirichlet state equal to each node's own
    state, i.e. the RHS of the same state with per-node Dirichlet values."""
    c = box(periodic=(0, 1))
    x = c.coordinates()
    u = np.zeros((c.E, c.ndofs, 4))
    u[..., 0] = 1.0
    u[..., 3] = (1.0 + 0.2 * np.cos(x[..., 1])) / (GAMMA - 1)
    for face in (0, 1):
        c.set_boundary(face, O.BC_SLIP_WALL)
    du_wall = c.rhs(u)
    # per-node Dirichlet states equal to the inner traces give the same fluxes
    d = box(periodic=(0, 1))
    n = 4
    for face, ie in ((0, 0), (1, 5)):
        vals = np.zeros((5 * n, 4))
        for je in range(5):
            e = ie * 5 + je
            for q in range(n):
                dof = q * n + (0 if face == 0 else n - 1)
                vals[je * n + q] = u[e, dof]
        d.set_boundary(face, O.BC_DIRICHLET, values=vals)
    du_dir = d.rhs(u)
    assert np.max(np.abs(du_wall - du_dir)) <= 1e-13 * np.max(np.abs(du_wall))


def test_per_node_dirichlet_equals_constant_dirichlet():
    c1, c2 = box(), box()
    u = random_state(c1, seed=3)
    state = np.array([1.1, 0.2, -0.1, 2.6])
    for face in range(4):
        c1.set_boundary(face, O.BC_DIRICHLET, state=state)
        count = 5 if face < 2 else 6
        c2.set_boundary(face, O.BC_DIRICHLET, values=np.tile(state, (count * 4, 1)))
    assert np.array_equal(c1.rhs(u), c2.rhs(u))
