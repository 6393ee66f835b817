"""GPU parity of the Navier-Stokes right-hand side (BR1 gradient lift,
Newtonian viscous flux, one-sided viscous interface flux) against the CPU
restatement.  The viscous contribution is small next to the inviscid one at
the TGV viscosity, so it is also compared on its own: (NS RHS - Euler RHS)
from the GPU against the same difference from the oracle."""
import math

import numpy as np
import pytest

import dgsem_amd as dg
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TWO_PI = 2 * math.pi


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def pair(dim, N, cells, left, right, mu, prandtl, eq, geometry=dg.GEOMETRY_NODAL):
    oc = O.Case(dim, N, cells, left, right, eq=eq, mu=mu, prandtl=prandtl)
    cfg = dg.SolverConfig(dim=dim, polydeg=N, cells=cells, domain_left=left, domain_right=right,
                          equation=eq, mu=mu, prandtl=prandtl, geometry=geometry)
    gs = dg.StructuredSolver(cfg)
    if geometry == dg.GEOMETRY_AFFINE:
        h = [(right[k] - left[k]) / cells[k] / 2.0 for k in range(dim)]
        Ja = np.zeros((oc.E, oc.ndofs, dim, dim))
        for r in range(dim):
            Ja[..., r, r] = np.prod([h[k] for k in range(dim) if k != r])
        oc.set_metrics(Ja, np.full((oc.E, oc.ndofs), 1.0 / np.prod(h)))
    return oc, gs


def gpu_rhs(gs, u_ref):
    u = gs.from_reference_layout(u_ref)
    du = gs.empty()
    gs.calc_rhs(u, du, 0.0)
    return gs.to_reference_layout(du)


@pytest.mark.parametrize("geometry", [dg.GEOMETRY_NODAL, dg.GEOMETRY_AFFINE])
@pytest.mark.parametrize("mu", [6.25e-4, 0.05])
def test_ns_tgv_rhs_and_viscous_part(gpu, mu, geometry):
    print(f"case ns tgv mu={mu} geometry={geometry}", flush=True)
    cells = (4, 3, 3)
    L, R = [0, 0, 0], [TWO_PI] * 3
    ons, gns = pair(3, 3, cells, L, R, mu, 0.72, dg.NAVIER_STOKES, geometry)
    oeu, geu = pair(3, 3, cells, L, R, mu, 0.72, dg.EULER, geometry)
    u = ons.init_tgv()
    ref_ns, got_ns = ons.rhs(u), gpu_rhs(gns, u)
    assert rel_l2(got_ns, ref_ns) <= 1e-12
    visc_ref = ref_ns - oeu.rhs(u)
    visc_got = got_ns - gpu_rhs(geu, u)
    assert rel_l2(visc_got, visc_ref) <= 1e-9


def test_ns_2d_khi_rhs(gpu):
    oc, gs = pair(2, 3, (8, 6), [-1, -1], [1, 1], 5e-3, 0.73, dg.NAVIER_STOKES)
    oe, ge = pair(2, 3, (8, 6), [-1, -1], [1, 1], 5e-3, 0.73, dg.EULER)
    u = oc.init_khi()
    ref, got = oc.rhs(u), gpu_rhs(gs, u)
    assert rel_l2(got, ref) <= 1e-12
    assert rel_l2(got - gpu_rhs(ge, u), ref - oe.rhs(u)) <= 1e-9


def test_ns_tgv_ssprk3_100_steps(gpu):
    oc, gs = pair(3, 3, (4, 4, 4), [0, 0, 0], [TWO_PI] * 3, 6.25e-4, 0.72, dg.NAVIER_STOKES)
    u0 = oc.init_tgv()
    dt = 1e-3
    ref = oc.ssprk3(u0, dt, 100)
    u = gs.from_reference_layout(u0)
    t1, t2 = gs.empty(), gs.empty()
    gs.ssprk3_steps(u, dt, 100, t1, t2)
    assert rel_l2(gs.to_reference_layout(u), ref) <= 1e-10
