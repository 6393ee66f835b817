"""Tangent and discrete adjoint of the right-hand side and of SSPRK3 steps.

Checks, each on the GPU kernels through the C ABI:
  * tangent vs central finite differences of the forward RHS (h = 1e-6);
  * adjoint vs tangent: <w, J v> = <J^T w, v> for random v, w (relative
    1e-12), for the RHS and for a full SSPRK3 step;
  * gradient of the time-integrated kinetic energy with respect to the
    initial velocity field, from the checkpointed reverse sweep and the
    torch.autograd.Function, against central finite differences (h = 1e-6)
    of the oracle's CPU solver along seeded Gaussian directions (relative
    1e-6).
"""
import math

import numpy as np
import pytest
import torch

import dgsem_amd as dg
from dgsem_amd import adjoint, tgv
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi


def solver(dim, N, cells, eq, geometry=dg.GEOMETRY_NODAL, mu=6.25e-4):
    left, right = ([0.0] * dim, [TWO_PI] * dim) if dim == 3 else ([-1.0, -1.0], [1.0, 1.0])
    cfg = dg.SolverConfig(dim=dim, polydeg=N, cells=cells, domain_left=left, domain_right=right,
                          equation=eq, mu=mu if eq == dg.NAVIER_STOKES else None,
                          prandtl=0.72 if dim == 3 else 0.73, geometry=geometry)
    return dg.StructuredSolver(cfg)


def initial_state(s):
    x = s.node_coordinates()
    if s.cfg.dim == 3:
        return s.from_reference_layout(tgv.tgv_state(x))
    from dgsem_amd import ensemble
    return s.from_reference_layout(ensemble.khi_state(x))


CASES = [
    (3, 3, (3, 3, 3), dg.EULER, dg.GEOMETRY_NODAL),
    (3, 3, (3, 3, 3), dg.NAVIER_STOKES, dg.GEOMETRY_NODAL),
    (3, 2, (3, 2, 4), dg.NAVIER_STOKES, dg.GEOMETRY_AFFINE),
    (3, 5, (2, 2, 2), dg.EULER, dg.GEOMETRY_AFFINE),
    (3, 7, (2, 1, 2), dg.EULER, dg.GEOMETRY_NODAL),
    (3, 7, (2, 1, 2), dg.NAVIER_STOKES, dg.GEOMETRY_NODAL),
    (2, 7, (3, 2), dg.NAVIER_STOKES, dg.GEOMETRY_AFFINE),
    (2, 3, (6, 5), dg.EULER, dg.GEOMETRY_NODAL),
    (2, 4, (4, 4), dg.NAVIER_STOKES, dg.GEOMETRY_AFFINE),
    (2, 1, (5, 4), dg.NAVIER_STOKES, dg.GEOMETRY_NODAL),
]


def ids(c):
    return f"{c[0]}d-N{c[1]}-{'ns' if c[3] else 'euler'}-{'nodal' if c[4] == 0 else 'affine'}"


@pytest.mark.parametrize("case", CASES, ids=[ids(c) for c in CASES])
def test_rhs_tangent_matches_finite_differences(gpu, case):
    dim, N, cells, eq, geo = case
    print(f"case jvp fd {ids(case)}", flush=True)
    s = solver(dim, N, cells, eq, geo, mu=0.05)
    u = initial_state(s)
    g = torch.Generator(device="cuda").manual_seed(3)
    v = torch.randn(s.shape, dtype=torch.float64, device="cuda", generator=g)
    jv = s.rhs_jvp(u, v)
    h = 1e-6
    lp, lm = s.empty(), s.empty()
    s.calc_rhs((u + h * v).contiguous(), lp)
    s.calc_rhs((u - h * v).contiguous(), lm)
    fd = (lp - lm) / (2 * h)
    err = (torch.linalg.norm(fd - jv) / torch.linalg.norm(jv)).item()
    assert err <= 1e-6, err


@pytest.mark.parametrize("case", CASES, ids=[ids(c) for c in CASES])
def test_rhs_adjoint_dot_product(gpu, case):
    dim, N, cells, eq, geo = case
    print(f"case vjp dot {ids(case)}", flush=True)
    s = solver(dim, N, cells, eq, geo, mu=0.05)
    u = initial_state(s)
    g = torch.Generator(device="cuda").manual_seed(5)
    v = torch.randn(s.shape, dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn(s.shape, dtype=torch.float64, device="cuda", generator=g)
    lhs = torch.sum(w * s.rhs_jvp(u, v)).item()
    rhs = torch.sum(s.rhs_vjp(u, w) * v).item()
    scale = torch.sum(torch.abs(w * s.rhs_jvp(u, v))).item()
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)


@pytest.mark.parametrize("eq", [dg.EULER, dg.NAVIER_STOKES])
def test_step_tangent_and_adjoint(gpu, eq):
    print(f"case step jvp/vjp eq={eq}", flush=True)
    s = solver(3, 3, (3, 3, 3), eq, mu=0.05)
    u0 = initial_state(s)
    dt = 2e-3
    u1, u2 = s.ssprk3_stages(u0, dt)
    g = torch.Generator(device="cuda").manual_seed(9)
    v = torch.randn(s.shape, dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn(s.shape, dtype=torch.float64, device="cuda", generator=g)
    jv = s.step_jvp(dt, u0, u1, u2, v)
    # tangent vs finite differences of one step
    h = 1e-6
    outs = []
    for sign in (1, -1):
        x = (u0 + sign * h * v).contiguous()
        t1, t2 = s.empty(), s.empty()
        s.ssprk3_steps(x, dt, 1, t1, t2)
        outs.append(x)
    fd = (outs[0] - outs[1]) / (2 * h)
    # the state carries E ~ 180, so rounding limits the difference quotient to ~1e-7
    assert (torch.linalg.norm(fd - jv) / torch.linalg.norm(jv)).item() <= 1e-5
    jtw = s.step_vjp(dt, u0, u1, u2, w)
    lhs, rhs = torch.sum(w * jv).item(), torch.sum(jtw * v).item()
    assert abs(lhs - rhs) <= 1e-12 * torch.sum(torch.abs(w * jv)).item()


def test_kinetic_energy_matches_oracle(gpu):
    s = solver(3, 3, (3, 3, 3), dg.EULER)
    oc = O.Case(3, 3, (3, 3, 3), [0, 0, 0], [TWO_PI] * 3)
    u = oc.init_tgv()
    ke = s.kinetic_energy(s.from_reference_layout(u)).item()
    assert abs(ke - oc.kinetic_energy(u)) <= 1e-13 * abs(ke)


def velocity_field(s):
    rho, v, p = tgv.tgv_primitives(s.node_coordinates())
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    return t(rho), t(np.transpose(v, (0, 2, 1))), t(p)  # v: [E][3][nn]


def oracle_objective(oc, rho, v, p, dt, nsteps):
    u = adjoint.conservative_from_velocity(rho, v, p)
    u = u.permute(0, 2, 1).contiguous().cpu().numpy()
    J = 0.0
    vals = []
    for _ in range(nsteps):
        vals.append(oc.kinetic_energy(u))
        u = oc.ssprk3(u, dt, 1)
    J = dt * math.fsum(vals)
    return J


def test_objective_gradient_matches_oracle_finite_differences(gpu):
    """3D Navier-Stokes TGV, 3^3 elements, N = 3, 12 steps, checkpoints every
    5 steps; 4 seeded Gaussian directions of the initial velocity field."""
    print("case objective gradient vs oracle fd", flush=True)
    cells, N, nsteps, every = (3, 3, 3), 3, 12, 5
    mu = 6.25e-4
    s = solver(3, N, cells, dg.NAVIER_STOKES, mu=mu)
    oc = O.Case(3, N, cells, [0, 0, 0], [TWO_PI] * 3, eq=O.NAVIER_STOKES, mu=mu, prandtl=0.72)
    dt = tgv.tgv_dt(cells[0], N)
    rho, v, p = velocity_field(s)
    vv = v.clone().requires_grad_(True)
    u0 = adjoint.conservative_from_velocity(rho, vv, p)
    J = adjoint.KineticEnergyObjective.apply(u0, s, dt, nsteps, every)
    J.backward()
    grad = vv.grad
    J_oracle = oracle_objective(oc, rho, v, p, dt, nsteps)
    assert abs(J.item() - J_oracle) <= 1e-10 * abs(J_oracle)
    rng = torch.Generator(device="cuda").manual_seed(2024)
    h = 1e-6
    for k in range(4):
        d = torch.randn(v.shape, dtype=torch.float64, device="cuda", generator=rng)
        fd = (oracle_objective(oc, rho, v + h * d, p, dt, nsteps) -
              oracle_objective(oc, rho, v - h * d, p, dt, nsteps)) / (2 * h)
        ad = torch.sum(grad * d).item()
        print(f"  direction {k}: adjoint {ad:.12e} fd {fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-6 * abs(fd)


def test_objective_gradient_full_size(gpu):
    """16^3 elements, N = 3, Navier-Stokes TGV at Re = 1600, 200 steps,
    checkpoints every 20 steps: the adjoint gradient against central finite
    differences of the GPU forward along 16 seeded Gaussian directions."""
    print("case objective gradient full size", flush=True)
    cells, N, nsteps, every = (16, 16, 16), 3, 200, 20
    s = solver(3, N, cells, dg.NAVIER_STOKES, mu=6.25e-4)
    dt = tgv.tgv_dt(cells[0], N)
    rho, v, p = velocity_field(s)
    vv = v.clone().requires_grad_(True)
    u0 = adjoint.conservative_from_velocity(rho, vv, p)
    J = adjoint.KineticEnergyObjective.apply(u0, s, dt, nsteps, every)
    J.backward()
    grad = vv.grad
    rng = torch.Generator(device="cuda").manual_seed(1600)
    h = 1e-6
    worst = 0.0
    for k in range(16):
        d = torch.randn(v.shape, dtype=torch.float64, device="cuda", generator=rng)
        Jp, _ = adjoint.kinetic_energy_trajectory(
            s, adjoint.conservative_from_velocity(rho, v + h * d, p), dt, nsteps)
        Jm, _ = adjoint.kinetic_energy_trajectory(
            s, adjoint.conservative_from_velocity(rho, v - h * d, p), dt, nsteps)
        fd = (Jp - Jm) / (2 * h)
        ad = torch.sum(grad * d).item()
        worst = max(worst, abs(ad - fd) / abs(fd))
    print(f"  worst relative difference {worst:.3e}", flush=True)
    assert worst <= 1e-6
