"""GPU parity of the Navier-Stokes adiabatic walls (no slip and slip) and of
the other boundary kinds under Navier-Stokes, against the oracle (pinned in
test_oracle_ns_walls.py).  The viscous contribution is also compared on its
own, (NS RHS - Euler RHS), with a viscosity large enough that the wall terms
carry weight.  The Daru-Tenaud viscous shock tube (examples/NavierStokes2D/
viscous_shock_tube: no-slip walls on three sides, a slip wall on top, shock
capturing with alpha_max = 1, alpha_min = 0.1) runs 100 SSPRK3 steps.

Bars: one RHS relative L2 <= 1e-12, the viscous part <= 1e-9, 100 steps
<= 1e-10, identical troubled sets."""
import numpy as np
import pytest

import dgsem_amd as dg
from gpu_pairs import gpu_rhs, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GAMMA, PR = 1.4, 0.73
NOSLIP, SLIP = dg.BC_NS_NOSLIP_WALL, dg.BC_NS_SLIP_WALL
LAYOUTS = {
    "noslip_box": [NOSLIP] * 4,
    "shock_tube": [NOSLIP, NOSLIP, NOSLIP, SLIP],
    "slip_channel": [SLIP, SLIP, dg.BC_PERIODIC, dg.BC_PERIODIC],
    "mixed_channel": [dg.BC_PERIODIC, dg.BC_PERIODIC, NOSLIP, SLIP],
    "euler_kinds": [dg.BC_DIRICHLET, dg.BC_OUTFLOW, dg.BC_SLIP_WALL, NOSLIP],
}
STATE = [1.3, 0.2, -0.1, 2.4]


def make_ns_pair(N, cells, faces, *, eq=dg.NAVIER_STOKES, mu=5e-3, geometry=dg.GEOMETRY_NODAL,
                 volume=dg.SPLIT_FORM, left=(-1.0, -0.5), right=(1.0, 0.75), amax=0.5, amin=0.001,
                 surface=dg.FLUX_LLF, fv=dg.FLUX_LLF):
    periodic = [int(faces[0] == dg.BC_PERIODIC), int(faces[2] == dg.BC_PERIODIC)]
    if eq == dg.EULER:  # the Euler twin: walls become slip walls
        faces = [dg.BC_SLIP_WALL if k in (NOSLIP, SLIP) else k for k in faces]
    oc = O.Case(2, N, cells, list(left), list(right), periodic=periodic, eq=eq, mu=mu,
                prandtl=PR, volume=volume, alpha_max=amax, alpha_min=amin, surface_flux=surface,
                fv_flux=fv)
    states = [STATE if k == dg.BC_DIRICHLET else None for k in faces]
    for f, kind in enumerate(faces):
        if kind != dg.BC_PERIODIC:
            oc.set_boundary(f, kind, state=states[f])
    cfg = dg.SolverConfig(dim=2, polydeg=N, cells=cells, domain_left=list(left),
                          domain_right=list(right), equation=eq, mu=mu, prandtl=PR,
                          volume_integral=volume, alpha_max=amax, alpha_min=amin,
                          surface_flux=surface, fv_flux=fv, geometry=geometry, face_boundary=faces,
                          face_state=[s or [0.0] * 4 for s in states])
    gs = dg.StructuredSolver(cfg)
    if geometry == dg.GEOMETRY_AFFINE:
        h = [(right[k] - left[k]) / cells[k] / 2.0 for k in range(2)]
        Ja = np.zeros((oc.E, oc.ndofs, 2, 2))
        Ja[..., 0, 0], Ja[..., 1, 1] = h[1], h[0]
        oc.set_metrics(Ja, np.full((oc.E, oc.ndofs), 1.0 / (h[0] * h[1])))
    return oc, gs


def smooth_state(oc, seed=0):
    rng = np.random.default_rng(seed)
    x = oc.coordinates()
    rho = 1.0 + 0.2 * np.sin(3 * x[..., 0]) * np.cos(2 * x[..., 1]) + 0.01 * rng.random(x.shape[:-1])
    vx = 0.3 * np.cos(2 * x[..., 1]) + 0.05 * rng.random(rho.shape)
    vy = 0.2 * np.sin(3 * x[..., 0])
    p = 1.0 + 0.1 * np.cos(x[..., 0] + x[..., 1])
    return np.stack([rho, rho * vx, rho * vy, p / (GAMMA - 1) + 0.5 * rho * (vx * vx + vy * vy)], -1)


def daru_tenaud(oc):
    """rho = 120 for x <= 0.5 and 1.2 beyond, at rest, p = rho / gamma."""
    x = oc.coordinates()
    rho = np.where(x[..., 0] <= 0.5, 120.0, 1.2)
    u = np.zeros(x.shape[:-1] + (4,))
    u[..., 0] = rho
    u[..., 3] = rho / GAMMA / (GAMMA - 1)
    return u


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("N,cells", [(1, (7, 6)), (2, (6, 5)), (3, (6, 5)), (4, (5, 4)),
                                     (5, (4, 4)), (7, (3, 3))])
def test_ns_wall_rhs(gpu, N, cells, layout):
    print(f"case ns wall rhs N={N} cells={cells} layout={layout}", flush=True)
    for geometry in (dg.GEOMETRY_NODAL, dg.GEOMETRY_AFFINE):
        oc, gs = make_ns_pair(N, cells, LAYOUTS[layout], geometry=geometry)
        u = smooth_state(oc, seed=N)
        assert rel_l2(gpu_rhs(gs, u), oc.rhs(u)) <= 1e-12, geometry


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_ns_wall_viscous_part(gpu, layout):
    print(f"case ns wall viscous part layout={layout}", flush=True)
    N, cells, mu = 3, (6, 5), 0.05
    ons, gns = make_ns_pair(N, cells, LAYOUTS[layout], mu=mu)
    oeu, geu = make_ns_pair(N, cells, LAYOUTS[layout], eq=dg.EULER, mu=mu)
    u = smooth_state(ons, seed=11)
    ref_ns, got_ns = ons.rhs(u), gpu_rhs(gns, u)
    assert rel_l2(got_ns, ref_ns) <= 1e-12
    visc_ref = ref_ns - oeu.rhs(u)
    visc_got = got_ns - gpu_rhs(geu, u)
    assert rel_l2(visc_got, visc_ref) <= 1e-9


def test_viscous_shock_tube_rhs_and_steps(gpu):
    print("case ns viscous shock tube", flush=True)
    N, cells, steps = 3, (16, 8), 100
    oc, gs = make_ns_pair(N, cells, LAYOUTS["shock_tube"], mu=1e-3, volume=dg.SHOCK_CAPTURING,
                          left=(0.0, 0.0), right=(1.0, 0.5), amax=1.0, amin=0.1)
    u0 = daru_tenaud(oc)
    ref, alpha_ref = oc.rhs(u0, with_alpha=True)
    got = gpu_rhs(gs, u0)
    alpha = gs.alpha().cpu().numpy()
    assert (alpha_ref > 0).any() and not (alpha_ref > 0).all()
    assert np.array_equal(alpha > 0, alpha_ref > 0)
    assert rel_l2(got, ref) <= 1e-12
    # dt as the example sets it: CFL 0.4, max wave speed 7
    dt = 0.4 * min(1.0 / cells[0], 0.5 / cells[1]) / ((2 * N + 1) * 7.0)
    ref = oc.ssprk3(u0, dt, steps)
    u = gs.from_reference_layout(u0)
    t1, t2 = gs.empty(), gs.empty()
    gs.ssprk3_steps(u, dt, steps, t1, t2)
    got = gs.to_reference_layout(u)
    assert np.isfinite(got).all()
    assert rel_l2(got, ref) <= 1e-10
    # closed box of adiabatic walls: mass and energy stay put
    w = O.basis(N)["weights"]
    vol = np.einsum("j,i->ji", w, w).ravel()[None, :] / np.abs(oc.inverse_jacobian())
    for var in (0, 3):
        t0, t = np.sum(vol * u0[..., var]), np.sum(vol * got[..., var])
        assert abs(t - t0) <= 1e-12 * abs(t0), var


def test_ns_dirichlet_heat_conduction(gpu):
    """Per-node Dirichlet states under Navier-Stokes: gas at rest with a
    linear temperature between the x faces (steady conduction); the GPU's
    gradient and viscous face values match the oracle's."""
    from test_oracle_ns_walls import face_traces
    print("case ns dirichlet conduction", flush=True)
    N, cells = 4, (4, 3)
    faces = [dg.BC_DIRICHLET, dg.BC_DIRICHLET, dg.BC_PERIODIC, dg.BC_PERIODIC]
    oc, gs = make_ns_pair(N, cells, faces, mu=0.05)
    x = oc.coordinates()
    T = 1.0 + 0.2 * x[..., 0] + 0.05 * np.sin(2 * x[..., 1])
    u = np.zeros(x.shape[:-1] + (4,))
    u[..., 0] = 1.0 / T
    u[..., 3] = 1.0 / (GAMMA - 1)
    keep = []
    for f in (0, 1):
        vals = face_traces(oc, u, f, N, cells) * 1.01  # off the inner traces
        oc.set_boundary(f, dg.BC_DIRICHLET, values=vals)
        keep.append(torch_tensor(vals))
        gs.set_boundary_values(f, keep[-1])
    assert rel_l2(gpu_rhs(gs, u), oc.rhs(u)) <= 1e-12


def torch_tensor(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()
