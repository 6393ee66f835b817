"""The oracle's positivity floor (a build extension applied after every
SSPRK3 stage update; the reference has none) pinned by its invariants:

  * density: every node ends at or above the floor, and a node below it ends
    exactly at the floor;
  * pressure: every node ends at or above the floor, up to the rounding of
    (gamma - 1) (E - kin) at the raised energy, and a raised node's energy is
    exactly floor / (gamma - 1) + kin with kin from the clamped density;
  * momentum is never touched;
  * a node that already satisfies both bounds (including one sitting exactly
    on the floor) is returned bit for bit.

A floor that clamped with <=, touched the momentum, raised the pressure with
the unclamped density or skipped a kind of node fails one of these.
"""
import numpy as np
import pytest

from oracle import oracle as O

FLOOR = 1e-13
GAMMA = 1.4


def pressure(u, gamma=GAMMA):
    d = u.shape[-1] - 2
    kin = 0.5 * np.sum(u[..., 1:1 + d] ** 2, axis=-1) / u[..., 0]
    return (gamma - 1.0) * (u[..., -1] - kin)


def planted_state(oc, rng):
    """A state whose nodes cover every branch: compliant, density below the
    floor, pressure below it, both, and exactly on the floor."""
    E, nn, nv = oc.shape
    d = nv - 2
    rho = rng.uniform(0.5, 2.0, (E, nn))
    v = rng.normal(0.0, 0.3, (E, nn, d))
    p = rng.uniform(0.5, 2.0, (E, nn))
    kind = rng.integers(0, 5, (E, nn))
    rho[kind == 1] = rng.uniform(-1e-13, 0.9e-13, np.count_nonzero(kind == 1))
    p[kind == 2] = rng.uniform(-1e-12, 0.9e-13, np.count_nonzero(kind == 2))
    both = kind == 3
    rho[both] = rng.uniform(1e-15, 0.9e-13, np.count_nonzero(both))
    p[both] = rng.uniform(1e-16, 0.9e-13, np.count_nonzero(both))
    rho[kind == 4] = FLOOR
    u = np.zeros((E, nn, nv))
    u[..., 0] = rho
    u[..., 1:1 + d] = rho[..., None] * v
    u[..., -1] = p / (GAMMA - 1.0) + 0.5 * rho * np.sum(v * v, axis=-1)
    # kinds 1-3 with density at or below zero: keep the momentum small so the
    # clamped state stays finite
    low = (kind == 1) | both
    u[low, 1:1 + d] = rng.normal(0.0, 1e-14, (np.count_nonzero(low), d))
    return u, kind


@pytest.mark.parametrize("dim", [2, 3])
def test_floor_invariants(dim):
    rng = np.random.default_rng(11 + dim)
    oc = O.Case(dim, 3, (3,) * dim, [-1] * dim, [1] * dim, positivity_floor=FLOOR)
    u, kind = planted_state(oc, rng)
    out = oc.apply_floor(u)
    d = dim
    # momentum never changes
    assert np.array_equal(out[..., 1:1 + d], u[..., 1:1 + d])
    # density
    assert np.all(out[..., 0] >= FLOOR)
    below = u[..., 0] < FLOOR
    assert np.any(below)
    assert np.all(out[below, 0] == FLOOR)
    assert np.array_equal(out[~below, 0], u[~below, 0])
    # pressure, with the clamped density
    kin = 0.5 * np.sum(out[..., 1:1 + d] ** 2, axis=-1) / out[..., 0]
    p_in = (GAMMA - 1.0) * (u[..., -1] - kin)
    raised = p_in < FLOOR
    assert np.any(raised) and np.any(~raised)
    assert np.all(out[raised, -1] == FLOOR / (GAMMA - 1.0) + kin[raised])
    assert np.array_equal(out[~raised, -1], u[~raised, -1])
    p_out = pressure(out)
    slack = 2.0 * (GAMMA - 1.0) * np.spacing(np.abs(out[..., -1]))
    assert np.all(p_out >= FLOOR - slack)
    # compliant nodes, including those exactly on the density floor: bitwise
    ok = ~below & ~raised
    assert np.any(ok & (kind == 4))
    assert np.array_equal(out[ok], u[ok])


def test_floor_disabled_is_identity():
    rng = np.random.default_rng(3)
    oc = O.Case(2, 3, (3, 3), [-1, -1], [1, 1], positivity_floor=0.0)
    u, _ = planted_state(oc, rng)
    assert np.array_equal(oc.apply_floor(u), u)


def slab_state(oc):
    """Still gas in three slabs along x (uniform pressure inside each, zero
    velocity): x < -1/3 rho = 1, p = 1; the middle slab rho = 5e-14, p = 1;
    x > 1/3 rho = 5e-14, p = 5e-14.  Away from the slab faces the mass and
    energy fluxes vanish, so stage updates leave density (and in the last
    slab pressure) below the floor there, and the floor's density and
    pressure branches both act."""
    x = oc.coordinates()[..., 0]
    xc = x.mean(axis=1, keepdims=True) + 0.0 * x  # slab by element centre
    rho = np.where(xc < -1.0 / 3.0, 1.0, 5e-14)
    p = np.where(xc > 1.0 / 3.0, 5e-14, 1.0)
    u = np.zeros(oc.shape)
    u[..., 0] = rho
    u[..., -1] = p / (GAMMA - 1.0)
    return u


@pytest.mark.parametrize("dim", [2, 3])
def test_floor_acts_in_ssprk3_steps(dim):
    """Both branches fire inside the oracle's SSPRK3 stages, and the step
    ends within the bounds."""
    cells = (6, 2) if dim == 2 else (6, 2, 2)
    oc = O.Case(dim, 3, cells, [-1] * dim, [1] * dim, positivity_floor=FLOOR)
    u0 = slab_state(oc)
    free = O.Case(dim, 3, cells, [-1] * dim, [1] * dim, positivity_floor=0.0)
    dt = 1e-9
    u1 = oc.ssprk3(u0, dt, 1)
    raw = free.ssprk3(u0, dt, 1)
    assert np.any(raw[..., 0] < FLOOR) and np.any(pressure(raw) < FLOOR)
    assert np.all(u1[..., 0] >= FLOOR)
    slack = 2.0 * (GAMMA - 1.0) * np.spacing(np.abs(u1[..., -1]))
    assert np.all(pressure(u1) >= FLOOR - slack)
    assert np.all(np.isfinite(u1))
