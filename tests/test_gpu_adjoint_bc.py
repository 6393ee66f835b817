"""Tangent and adjoint of the right-hand side on bounded 2D meshes: slip
walls, outflow, Dirichlet states (constant and per node) and the mixed
Dirichlet/slip-wall face of the double Mach reflection, with the split form
or shock capturing (alpha frozen), LLF or HLLC; and Navier-Stokes with its
adiabatic no-slip and slip walls.  At a non-periodic face the flux takes the
element's own trace and the outer state the boundary condition forms from it
(boundary_set.hpp:88-111 order of faces): a slip wall mirrors the momentum
about the face, outflow takes the mean of the line's nodes, a Dirichlet state
is constant; the derivatives carry the outer state's dependence on the inner
one.  Under Navier-Stokes the BR1 gradient's face value at a wall is the
wall map of the own node's variables (navier_stokes_wall_boundary.hpp), and
the viscous face flux is the own node's, for the variables the condition
carries.

The states drift (|v.n| > 0 on every face node, and not equal to the
Dirichlet inflow's), and both velocity components carry node-wise noise: the
two traces on an interior face sit at the same point, so a smooth velocity
field alone would give them the same v.n, and the LLF wave speed's max()
would tie there, a kink that a central difference straddles.  The only ties
left are the ones a slip wall makes: there the mirror state has the same
|v.n| and sound speed as the inner one for every state, so both branches of
the max() have the same derivative.

Bars: tangent against central differences (h = 1e-7) of the oracle's RHS,
alpha overridden to the frozen values, relative 1e-6; <w, J v> = <J^T w, v>
to 1e-12 relative, for the RHS and for one SSPRK3 step."""
import numpy as np
import pytest
import torch

import dgsem_amd as dg
from oracle import oracle as O
from test_gpu_adjoint_sc import direction
from test_gpu_boundaries import INFLOW, LAYOUTS
from test_gpu_ns_walls import LAYOUTS as NS_LAYOUTS
from test_gpu_ns_walls import make_ns_pair

pytestmark = pytest.mark.gpu

GAMMA = 1.4
LEFT, RIGHT = (-1.0, -0.5), (1.0, 0.75)
H, L = dg.FLUX_HLLC, dg.FLUX_LLF

CASES = [
    # layout, volume integral, surface flux, subcell flux, geometry, N, cells
    ("walls", dg.SPLIT_FORM, L, L, dg.GEOMETRY_NODAL, 3, (6, 5)),
    ("channel", dg.SPLIT_FORM, L, L, dg.GEOMETRY_AFFINE, 3, (6, 5)),
    ("channel", dg.SPLIT_FORM, H, L, dg.GEOMETRY_NODAL, 2, (6, 5)),
    ("x_walls", dg.SHOCK_CAPTURING, L, L, dg.GEOMETRY_NODAL, 3, (6, 5)),
    ("y_open", dg.SHOCK_CAPTURING, H, H, dg.GEOMETRY_NODAL, 4, (5, 4)),
    ("channel", dg.SHOCK_CAPTURING, L, H, dg.GEOMETRY_NODAL, 3, (6, 5)),
    ("mixed", dg.SHOCK_CAPTURING, L, L, dg.GEOMETRY_NODAL, 3, (8, 5)),
    ("mixed", dg.SPLIT_FORM, H, L, dg.GEOMETRY_AFFINE, 3, (8, 5)),
]

MIXED = [dg.BC_DIRICHLET, dg.BC_OUTFLOW, dg.BC_MIXED_DIRICHLET_SLIP_WALL, dg.BC_DIRICHLET]


def ids(c):
    name = {H: "hllc", L: "llf"}
    return (f"{c[0]}-{'sc' if c[1] == dg.SHOCK_CAPTURING else 'split'}-{name[c[2]]}-{name[c[3]]}-"
            f"{'nodal' if c[4] == dg.GEOMETRY_NODAL else 'affine'}-N{c[5]}")


def drifting_state(x, seed):
    """A smooth field in a uniform drift with a density and pressure jump
    across x = 0.1 (troubled and untroubled elements); x: [e][dof][2]."""
    rng = np.random.default_rng(seed)
    X, Y = x[..., 0], x[..., 1]
    rho = 1.0 + 0.2 * np.sin(3 * X) * np.cos(2 * Y) + 0.01 * rng.random(X.shape)
    p = 1.0 + 0.1 * np.cos(X + Y)
    rho = np.where(X < 0.1, rho + 0.8, rho)
    p = np.where(X < 0.1, p + 1.5, p)
    vx = 0.35 + 0.1 * np.cos(2 * Y) + 0.02 * rng.random(X.shape)
    vy = 0.25 + 0.15 * np.sin(3 * X) + 0.02 * rng.random(X.shape)
    return np.stack([rho, rho * vx, rho * vy, p / (GAMMA - 1) + 0.5 * rho * (vx * vx + vy * vy)], -1)


def pair(case):
    layout, volume, surface, fv, geometry, N, cells = case
    faces = MIXED if layout == "mixed" else LAYOUTS[layout]
    periodic = [int(faces[0] == dg.BC_PERIODIC), int(faces[2] == dg.BC_PERIODIC)]
    oc = O.Case(2, N, cells, list(LEFT), list(RIGHT), periodic=periodic, volume=volume,
                surface_flux=surface, fv_flux=fv, alpha_max=0.5, alpha_min=0.001)
    states = [INFLOW if k == dg.BC_DIRICHLET else None for k in faces]
    cfg = dg.SolverConfig(dim=2, polydeg=N, cells=cells, domain_left=list(LEFT),
                          domain_right=list(RIGHT), volume_integral=volume, surface_flux=surface,
                          fv_flux=fv, alpha_max=0.5, alpha_min=0.001, geometry=geometry,
                          face_boundary=faces, face_state=[s or [0.0] * 4 for s in states])
    gs = dg.StructuredSolver(cfg)
    for f, kind in enumerate(faces):
        if kind not in (dg.BC_PERIODIC, dg.BC_MIXED_DIRICHLET_SLIP_WALL):
            oc.set_boundary(f, kind, state=states[f])
    if layout == "mixed":
        n = N + 1
        x_nodes = np.linspace(LEFT[0], RIGHT[0], cells[0] * n)
        vals = np.column_stack([np.tile(INFLOW, (cells[0] * n, 1)), (x_nodes < -0.4).astype(float)])
        oc.set_boundary(2, O.BC_MIXED_DIRICHLET_SLIP_WALL, values=vals)
        gs.set_boundary_values(2, torch.from_numpy(vals).cuda())
    if geometry == dg.GEOMETRY_AFFINE:
        h = [(RIGHT[k] - LEFT[k]) / cells[k] / 2.0 for k in range(2)]
        Ja = np.zeros((oc.E, oc.ndofs, 2, 2))
        Ja[..., 0, 0], Ja[..., 1, 1] = h[1], h[0]
        oc.set_metrics(Ja, np.full((oc.E, oc.ndofs), 1.0 / (h[0] * h[1])))
    return oc, gs, drifting_state(oc.coordinates(), seed=N)


@pytest.mark.parametrize("case", CASES, ids=[ids(c) for c in CASES])
def test_tangent_matches_oracle_finite_differences(gpu, case):
    print(f"case bc jvp fd {ids(case)}", flush=True)
    oc, gs, u0 = pair(case)
    sc = case[1] == dg.SHOCK_CAPTURING
    if sc:
        _, alpha = oc.rhs(u0, with_alpha=True)
        assert 0 < int((alpha > 0).sum()) < oc.E
        oc.set_alpha_override(alpha)
    v = direction(u0, 4)
    h = 1e-7
    fd = (oc.rhs(u0 + h * v) - oc.rhs(u0 - h * v)) / (2 * h)
    if sc:
        oc.set_alpha_override(None)
    u, vt = gs.from_reference_layout(u0), gs.from_reference_layout(v)
    jv = gs.to_reference_layout(gs.rhs_jvp(u, vt))
    err = float(np.linalg.norm(jv - fd) / np.linalg.norm(fd))
    print(f"  rel {err:.2e}", flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("case", CASES, ids=[ids(c) for c in CASES])
def test_adjoint_dot_product(gpu, case):
    print(f"case bc vjp dot {ids(case)}", flush=True)
    _, gs, u0 = pair(case)
    u = gs.from_reference_layout(u0)
    g = torch.Generator(device="cuda").manual_seed(7)
    v = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g)
    jv = gs.rhs_jvp(u, v)
    lhs = torch.sum(w * jv).item()
    rhs = torch.sum(gs.rhs_vjp(u, w) * v).item()
    scale = torch.sum(torch.abs(w * jv)).item()
    print(f"  {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale


@pytest.mark.parametrize("case", [CASES[1], CASES[6]], ids=[ids(CASES[1]), ids(CASES[6])])
def test_step_adjoint_dot_product(gpu, case):
    print(f"case bc step dot {ids(case)}", flush=True)
    _, gs, u0 = pair(case)
    u = gs.from_reference_layout(u0)
    dt = 1e-4
    u1, u2 = gs.ssprk3_stages(u, dt)
    g = torch.Generator(device="cuda").manual_seed(11)
    v = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g)
    jv = gs.step_jvp(dt, u, u1, u2, v)
    jtw = gs.step_vjp(dt, u, u1, u2, w)
    lhs, rhs = torch.sum(w * jv).item(), torch.sum(jtw * v).item()
    assert abs(lhs - rhs) <= 1e-12 * torch.sum(torch.abs(w * jv)).item()


NS_CASES = [
    # layout, volume integral, geometry, N, cells, surface and subcell flux
    ("noslip_box", dg.SPLIT_FORM, dg.GEOMETRY_NODAL, 3, (6, 5), L),
    ("shock_tube", dg.SHOCK_CAPTURING, dg.GEOMETRY_NODAL, 3, (6, 5), L),
    ("euler_kinds", dg.SPLIT_FORM, dg.GEOMETRY_AFFINE, 2, (6, 5), L),
    ("mixed_channel", dg.SPLIT_FORM, dg.GEOMETRY_NODAL, 4, (5, 4), L),
    ("slip_channel", dg.SHOCK_CAPTURING, dg.GEOMETRY_AFFINE, 3, (6, 5), L),
    ("shock_tube", dg.SPLIT_FORM, dg.GEOMETRY_NODAL, 7, (3, 3), L),
    ("shock_tube", dg.SHOCK_CAPTURING, dg.GEOMETRY_NODAL, 3, (6, 5), H),
    ("euler_kinds", dg.SPLIT_FORM, dg.GEOMETRY_AFFINE, 3, (6, 5), H),
]


def ns_ids(c):
    return (f"ns-{c[0]}-{'sc' if c[1] == dg.SHOCK_CAPTURING else 'split'}-"
            f"{'nodal' if c[2] == dg.GEOMETRY_NODAL else 'affine'}-N{c[3]}-{'hllc' if c[5] == H else 'llf'}")


def ns_pair(case):
    layout, volume, geometry, N, cells, flux = case
    oc, gs = make_ns_pair(N, cells, NS_LAYOUTS[layout], mu=0.02, geometry=geometry, volume=volume,
                          left=LEFT, right=RIGHT, surface=flux, fv=flux)
    return oc, gs, drifting_state(oc.coordinates(), seed=N)


@pytest.mark.parametrize("case", NS_CASES, ids=[ns_ids(c) for c in NS_CASES])
def test_navier_stokes_tangent_matches_oracle_finite_differences(gpu, case):
    print(f"case bc ns jvp fd {ns_ids(case)}", flush=True)
    oc, gs, u0 = ns_pair(case)
    sc = case[1] == dg.SHOCK_CAPTURING
    if sc:
        _, alpha = oc.rhs(u0, with_alpha=True)
        oc.set_alpha_override(alpha)
    v = direction(u0, 4)
    h = 1e-7
    fd = (oc.rhs(u0 + h * v) - oc.rhs(u0 - h * v)) / (2 * h)
    if sc:
        oc.set_alpha_override(None)
    u, vt = gs.from_reference_layout(u0), gs.from_reference_layout(v)
    jv = gs.to_reference_layout(gs.rhs_jvp(u, vt))
    err = float(np.linalg.norm(jv - fd) / np.linalg.norm(fd))
    print(f"  rel {err:.2e}", flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("case", NS_CASES, ids=[ns_ids(c) for c in NS_CASES])
def test_navier_stokes_adjoint_dot_product(gpu, case):
    print(f"case bc ns vjp dot {ns_ids(case)}", flush=True)
    _, gs, u0 = ns_pair(case)
    u = gs.from_reference_layout(u0)
    g = torch.Generator(device="cuda").manual_seed(7)
    v = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g)
    jv = gs.rhs_jvp(u, v)
    lhs = torch.sum(w * jv).item()
    rhs = torch.sum(gs.rhs_vjp(u, w) * v).item()
    scale = torch.sum(torch.abs(w * jv)).item()
    print(f"  {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale
    u1, u2 = gs.ssprk3_stages(u, 1e-4)
    jv = gs.step_jvp(1e-4, u, u1, u2, v)
    lhs, rhs = torch.sum(w * jv).item(), torch.sum(gs.step_vjp(1e-4, u, u1, u2, w) * v).item()
    assert abs(lhs - rhs) <= 1e-12 * torch.sum(torch.abs(w * jv)).item()


def test_navier_stokes_walls_objective_gradient(gpu):
    """The checkpointed objective gradient on a bounded mesh: the viscous
    shock tube's walls (no slip on three faces, slip on top), 10 SSPRK3
    steps, checkpoints every 4 steps, the kinetic-energy objective; the
    gradient with respect to the initial state against central differences
    of the GPU forward along 3 seeded Gaussian directions, relative 1e-6."""
    from dgsem_amd import adjoint
    print("case bc ns objective gradient shock_tube", flush=True)
    _, gs, u0 = ns_pair(("shock_tube", dg.SPLIT_FORM, dg.GEOMETRY_NODAL, 3, (6, 5), L))
    u = gs.from_reference_layout(u0)
    dt, nsteps, every = 1e-3, 10, 4
    uu = u.clone().requires_grad_(True)
    J = adjoint.KineticEnergyObjective.apply(uu, gs, dt, nsteps, every)
    J.backward()
    g = torch.Generator(device="cuda").manual_seed(5)
    h = 1e-6
    for k in range(3):
        d = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g)
        jp, _ = adjoint.kinetic_energy_trajectory(gs, u + h * d, dt, nsteps)
        jm, _ = adjoint.kinetic_energy_trajectory(gs, u - h * d, dt, nsteps)
        fd = (jp - jm) / (2 * h)
        ad = torch.sum(uu.grad * d).item()
        print(f"  direction {k}: adjoint {ad:.12e} fd {fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-6 * abs(fd)
