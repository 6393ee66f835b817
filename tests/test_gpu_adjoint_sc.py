"""Tangent and adjoint of the shock-capturing right-hand side with the
blending factor alpha frozen: the blend (1 - alpha) split form + alpha FV
(volume_flux.hpp:449-630, :821-869) with LLF, or HLLC, as both the surface
and the subcell flux, on the Kelvin-Helmholtz (2D) and Sedov (3D, LLF)
configurations and, for HLLC, the Kelvin-Helmholtz state in a uniform drift
(2D) and a Gaussian bump of density and pressure in a uniform flow (3D),
each with troubled and untroubled elements.  The drift keeps HLLC's
contact speed away from zero on every face, where the flux switches from
the left star state to the right one.

Checks:
  * the forward records each stage's alpha (dgsem_get_stage_alpha) and a
    frozen alpha (dgsem_freeze_alpha) reproduces the indicator's RHS bit for
    bit;
  * tangent against central finite differences (h = 1e-7) of the oracle's
    RHS with alpha overridden to the same frozen values (relative 1e-6);
  * adjoint against tangent: <w, J v> = <J^T w, v> (relative 1e-12), for the
    RHS and for one SSPRK3 step;
  * the checkpointed gradient of the time-integrated kinetic energy of a
    Kelvin-Helmholtz run against central finite differences of the GPU
    forward with every stage's alpha frozen at the unperturbed run's values
    (relative 1e-6);
  * with the positivity floor: the RHS derivatives unchanged by it, and the
    step derivatives through its density and pressure clamps (tangent
    against central differences of the GPU step, relative 1e-5; the
    dot-product identity).

The 2D states are the Kelvin-Helmholtz state with a smooth perturbation of
density, velocity and pressure.  The plain state is symmetric: along x its
x-velocity and sound speed are constant, so the LLF wave speed's two max()
arguments tie on every x-direction subcell face.  The sound speeds tie only
up to the rounding of the recomputed pressure, a kink some 1e-16 away, which
a central difference straddles and a derivative cannot see.  The
perturbation moves every such kink far from the state, and it is periodic
on the box: a jump across the periodic boundary in the x-velocity alone
would leave the sound speeds on both sides equal up to rounding, with a
nonzero jump for the wave speed to multiply.  No velocity component
vanishes at a node either, where |v.n| has its kink: the shear layer's
x-velocity is zero on the element faces y = +-0.5, and the x-velocity
perturbation is nonzero there.  The finite
differences use h = 1e-7, which keeps the O(h) error of an exact tie
(equal traces on element faces) below 2e-7.
"""
import numpy as np
import pytest
import torch

import dgsem_amd as dg
from dgsem_amd import adjoint, ensemble
from gpu_pairs import make_pair

pytestmark = pytest.mark.gpu

CASES = [
    # dim, N, cells, alpha_max, geometry, surface and subcell flux
    (2, 3, (8, 8), 0.002, dg.GEOMETRY_NODAL, dg.FLUX_LLF),
    (2, 3, (8, 8), 0.5, dg.GEOMETRY_AFFINE, dg.FLUX_LLF),
    (2, 4, (6, 5), 0.5, dg.GEOMETRY_NODAL, dg.FLUX_LLF),
    (3, 3, (4, 4, 4), 0.5, dg.GEOMETRY_NODAL, dg.FLUX_LLF),
    (3, 2, (4, 4, 4), 0.5, dg.GEOMETRY_AFFINE, dg.FLUX_LLF),
    (2, 3, (8, 8), 0.5, dg.GEOMETRY_NODAL, dg.FLUX_HLLC),
    (3, 3, (4, 4, 4), 0.5, dg.GEOMETRY_NODAL, dg.FLUX_HLLC),
]


def ids(c):
    return (f"{c[0]}d-N{c[1]}-amax{c[3]}-{'nodal' if c[4] == dg.GEOMETRY_NODAL else 'affine'}"
            f"-{'hllc' if c[5] == dg.FLUX_HLLC else 'llf'}")


def perturbed_khi(x, gamma=1.4, drift=(0.0, 0.0)):
    """The Kelvin-Helmholtz state (reference layout) with smooth
    perturbations of density, velocity and pressure and a uniform drift
    velocity added; x: [e][dof][2]."""
    X, Y = x[..., 0], x[..., 1]
    B = np.tanh(15 * Y + 7.5) - np.tanh(15 * Y - 7.5)
    rho = (0.5 + 0.75 * B) * (1 + 0.05 * np.sin(2 * np.pi * X + 0.7) * np.sin(np.pi * Y + 0.3))
    vx = 0.5 * (B - 1) + 0.03 * np.cos(np.pi * X + 0.4) * np.cos(np.pi * Y + 0.3) + drift[0]
    vy = 0.1 * np.sin(2 * np.pi * X) + 0.02 * np.sin(np.pi * Y + 0.9) + drift[1]
    p = 1.0 + 0.05 * np.cos(2 * np.pi * X + 1.1) * np.sin(2 * np.pi * Y + 0.2)
    E = p / (gamma - 1) + 0.5 * rho * (vx * vx + vy * vy)
    return np.stack([rho, rho * vx, rho * vy, E], axis=-1)


def drifting_bump(x, gamma=1.4):
    """A Gaussian bump of density and pressure at the centre of [-1, 1]^3
    in a uniform flow with smooth periodic perturbations; x: [e][dof][3]."""
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    g = np.exp(-(X * X + Y * Y + Z * Z) / 0.08)
    rho = 1 + 0.5 * g + 0.05 * np.sin(np.pi * X + 0.3) * np.sin(np.pi * Y + 0.5)
    vx = 0.5 + 0.05 * np.sin(np.pi * Y + 0.2)
    vy = 0.4 + 0.05 * np.sin(np.pi * Z + 0.6)
    vz = 0.3 + 0.05 * np.sin(np.pi * X + 1.0)
    p = 1 + 3 * g
    E = p / (gamma - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    return np.stack([rho, rho * vx, rho * vy, rho * vz, E], axis=-1)


def pair(case):
    dim, N, cells, amax, geo, flux = case
    box = ([-1.0] * dim, [1.0] * dim)
    oc, gs = make_pair(dim, N, cells, *box, geometry=geo, volume=dg.SHOCK_CAPTURING, surface=flux,
                       fv=flux, amax=amax, amin=0.001)
    x = oc.coordinates()
    if flux == dg.FLUX_HLLC:
        u = perturbed_khi(x, drift=(0.7, 0.3)) if dim == 2 else drifting_bump(x)
    else:
        u = perturbed_khi(x) if dim == 2 else oc.init_sedov()
    return oc, gs, u


def direction(u, seed):
    """A random direction scaled per variable like the state."""
    rng = np.random.default_rng(seed)
    scale = np.mean(np.abs(u), axis=(0, 1)) + 1e-3
    return rng.standard_normal(u.shape) * (np.abs(u) + scale)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", CASES, ids=[ids(c) for c in CASES])
def test_frozen_alpha_reproduces_the_indicator(gpu, case):
    oc, gs, u0 = pair(case)
    u = gs.from_reference_layout(u0)
    du, du2 = gs.empty(), gs.empty()
    gs.calc_rhs(u, du)
    alpha = gs.stage_alpha(0)
    assert torch.equal(alpha, gs.alpha())
    troubled = int((alpha > 0).sum())
    print(f"case {ids(case)}: {troubled} of {gs.n_elements} elements troubled", flush=True)
    assert 0 < troubled < gs.n_elements
    gs.freeze_alpha(alpha)
    gs.calc_rhs(u, du2)
    gs.freeze_alpha(None)
    assert torch.equal(du, du2)
    # the three stages of a step each keep their alpha
    t1, t2, t3 = gs.empty(), gs.empty(), gs.empty()
    gs.ssprk3_stage(1, 1e-4, None, u, t1)
    gs.ssprk3_stage(2, 1e-4, u, t1, t2)
    gs.ssprk3_stage(3, 1e-4, u, t2, t3)
    for k, x in ((1, u), (2, t1), (3, t2)):
        gs.calc_rhs(x, du)
        assert torch.equal(gs.stage_alpha(k), gs.stage_alpha(0)), k
    assert torch.equal(gs.alpha(), gs.stage_alpha(0))


@pytest.mark.parametrize("case", CASES, ids=[ids(c) for c in CASES])
def test_tangent_matches_oracle_finite_differences(gpu, case):
    oc, gs, u0 = pair(case)
    _, alpha = oc.rhs(u0, with_alpha=True)
    oc.set_alpha_override(alpha)
    v = direction(u0, 4)
    h = 1e-7
    fd = (oc.rhs(u0 + h * v) - oc.rhs(u0 - h * v)) / (2 * h)
    oc.set_alpha_override(None)
    u, vt = gs.from_reference_layout(u0), gs.from_reference_layout(v)
    jv = gs.to_reference_layout(gs.rhs_jvp(u, vt))
    err = rel(jv, fd)
    print(f"case jvp fd {ids(case)}: rel {err:.2e}", flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("case", CASES, ids=[ids(c) for c in CASES])
def test_adjoint_dot_product(gpu, case):
    _, gs, u0 = pair(case)
    u = gs.from_reference_layout(u0)
    g = torch.Generator(device="cuda").manual_seed(7)
    v = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g)
    jv = gs.rhs_jvp(u, v)
    lhs = torch.sum(w * jv).item()
    rhs = torch.sum(gs.rhs_vjp(u, w) * v).item()
    scale = torch.sum(torch.abs(w * jv)).item()
    print(f"case vjp dot {ids(case)}: {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale


STEP_CASES = [CASES[1], CASES[3], CASES[6]]


@pytest.mark.parametrize("case", STEP_CASES, ids=[ids(c) for c in STEP_CASES])
def test_step_adjoint_dot_product(gpu, case):
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


def frozen_objective(gs, u0, dt, nsteps, alphas):
    """J = sum_{n < nsteps} dt KE(u_n) of a forward run whose stages blend
    with the given alphas (alphas[n][k] for stage k + 1 of step n)."""
    ke = torch.empty(nsteps, dtype=torch.float64, device="cuda")
    u, u1, u2, nxt = u0.clone(), gs.empty(), gs.empty(), gs.empty()
    for n in range(nsteps):
        gs.kinetic_energy(u, ke[n:n + 1])
        gs.freeze_alpha(alphas[n][0])
        gs.ssprk3_stage(1, dt, None, u, u1)
        gs.freeze_alpha(alphas[n][1])
        gs.ssprk3_stage(2, dt, u, u1, u2)
        gs.freeze_alpha(alphas[n][2])
        gs.ssprk3_stage(3, dt, u, u2, nxt)
        u, nxt = nxt, u
    gs.freeze_alpha(None)
    return dt * float(np.sum(ke.cpu().numpy()))


@pytest.mark.parametrize("amax", [0.002, 0.5])
def test_checkpointed_khi_gradient_matches_frozen_finite_differences(gpu, amax):
    """Kelvin-Helmholtz 8^2 (perturbed), N = 3, 12 steps at the CFL-0.5
    step, checkpoints every 5 steps."""
    N, cells, nsteps, every = 3, (8, 8), 12, 5
    cfg = dg.SolverConfig(dim=2, polydeg=N, cells=cells, domain_left=(-1, -1), domain_right=(1, 1),
                          volume_integral=dg.SHOCK_CAPTURING, alpha_max=amax, alpha_min=0.001)
    gs = dg.StructuredSolver(cfg)
    x = gs.node_coordinates()
    ref = perturbed_khi(x)
    dt = ensemble.khi_dt(ref, N, cells)
    u0 = gs.from_reference_layout(ref)
    # the unperturbed run's alpha of every stage
    alphas = []
    u, u1, u2, nxt = u0.clone(), gs.empty(), gs.empty(), gs.empty()
    for _ in range(nsteps):
        gs.ssprk3_stage(1, dt, None, u, u1)
        gs.ssprk3_stage(2, dt, u, u1, u2)
        gs.ssprk3_stage(3, dt, u, u2, nxt)
        alphas.append([gs.stage_alpha(k).clone() for k in (1, 2, 3)])
        u, nxt = nxt, u
    assert any(bool((a > 0).any()) for step in alphas for a in step)
    J, grad = adjoint.kinetic_energy_objective(gs, u0, dt, nsteps, every)
    J0 = frozen_objective(gs, u0, dt, nsteps, alphas)
    assert abs(J - J0) <= 1e-14 * abs(J)
    g = torch.Generator(device="cuda").manual_seed(13)
    worst = 0.0
    for _ in range(3):
        d = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g) * u0.abs()
        h = 1e-7
        fd = (frozen_objective(gs, (u0 + h * d).contiguous(), dt, nsteps, alphas)
              - frozen_objective(gs, (u0 - h * d).contiguous(), dt, nsteps, alphas)) / (2 * h)
        ad = torch.sum(grad * d).item()
        worst = max(worst, abs(fd - ad) / abs(ad))
    print(f"case khi objective amax={amax}: J {J:.6e}, worst rel {worst:.2e}", flush=True)
    assert worst <= 1e-6


def test_positivity_floor_rhs(gpu):
    """The positivity floor clamps the stage results; the RHS is the same
    with or without it, and so are its tangent and adjoint."""
    kw = dict(dim=2, polydeg=3, cells=(4, 4), domain_left=(-1, -1), domain_right=(1, 1),
              volume_integral=dg.SHOCK_CAPTURING)
    plain = dg.StructuredSolver(dg.SolverConfig(**kw))
    gs = dg.StructuredSolver(dg.SolverConfig(positivity_floor=1e-10, **kw))
    u = gs.from_reference_layout(perturbed_khi(gs.node_coordinates()))
    g = torch.Generator(device="cuda").manual_seed(3)
    v = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g)
    assert torch.equal(gs.rhs_jvp(u, v), plain.rhs_jvp(u, v))
    assert torch.equal(gs.rhs_vjp(u, v), plain.rhs_vjp(u, v))


FLOOR_CASES = [
    # the pressure clamp: the drifting bump with its pressure scaled to
    # 0.1 .. 0.4 and the floor at 0.15 (density about 1); the density clamp:
    # the Kelvin-Helmholtz state (rho 0.5 .. 1.25) with the floor at 0.6.
    # (The Sedov state is at rest: |v.n| and the HLLC contact speed have
    # their kinks at it.)
    ("bump-pressure", 3, 3, (4, 4, 4), dg.FLUX_HLLC, 0.15),
    ("khi-density", 2, 3, (8, 8), dg.FLUX_LLF, 0.6),
]


def low_pressure_bump(x, gamma=1.4):
    """drifting_bump with the pressure 0.1 (1 + 3 g) in place of 1 + 3 g."""
    u = drifting_bump(x, gamma)
    kin = 0.5 * np.sum(u[..., 1:4] ** 2, axis=-1) / u[..., 0]
    p = (gamma - 1) * (u[..., 4] - kin)
    u[..., 4] = 0.1 * p / (gamma - 1) + kin
    return u


@pytest.mark.parametrize("case", FLOOR_CASES, ids=[c[0] for c in FLOOR_CASES])
def test_positivity_floor_step(gpu, case):
    """The step derivatives through the clamp of each stage's result: the
    tangent against central differences of the GPU step (h = 1e-7 along a
    direction scaled like the state at each node; the clamps act by a wide
    margin at their nodes), and the tangent against the adjoint by the
    dot-product identity."""
    name, dim, N, cells, fv, floor = case
    print(f"case floor step {name}", flush=True)
    box = ([-1.0] * dim, [1.0] * dim)
    kw = dict(dim=dim, polydeg=N, cells=cells, domain_left=box[0], domain_right=box[1],
              volume_integral=dg.SHOCK_CAPTURING, surface_flux=fv, fv_flux=fv, alpha_max=0.5,
              alpha_min=0.001)
    gs = dg.StructuredSolver(dg.SolverConfig(positivity_floor=floor, **kw))
    plain = dg.StructuredSolver(dg.SolverConfig(**kw))
    x = gs.node_coordinates()
    u0 = low_pressure_bump(x) if dim == 3 else perturbed_khi(x)
    dt = 2e-3
    u = gs.from_reference_layout(u0)
    # the clamp acts: one step with and without the floor differ
    a, b = u.clone(), u.clone()
    t1, t2 = gs.empty(), gs.empty()
    gs.ssprk3_steps(a, dt, 1, t1, t2)
    plain.ssprk3_steps(b, dt, 1, t1, t2)
    assert not torch.equal(a, b)
    u1, u2 = gs.ssprk3_stages(u, dt)
    # a direction scaled like the state node by node, so that the difference
    # quotient perturbs every node relatively (and no clamp margin is crossed)
    v = gs.from_reference_layout(direction(u0, 13))
    g = torch.Generator(device="cuda").manual_seed(13)
    w = torch.randn(gs.shape, dtype=torch.float64, device="cuda", generator=g)
    jv = gs.step_jvp(dt, u, u1, u2, v)
    h = 1e-7
    outs = []
    for sign in (1, -1):
        y = (u + sign * h * v).contiguous()
        gs.ssprk3_steps(y, dt, 1, t1, t2)
        outs.append(y)
    fd = (outs[0] - outs[1]) / (2 * h)
    err = (torch.linalg.norm(fd - jv) / torch.linalg.norm(jv)).item()
    print(f"  jvp vs fd rel {err:.2e}", flush=True)
    assert err <= 1e-5
    jtw = gs.step_vjp(dt, u, u1, u2, w)
    lhs, rhs = torch.sum(w * jv).item(), torch.sum(jtw * v).item()
    assert abs(lhs - rhs) <= 1e-12 * torch.sum(torch.abs(w * jv)).item()
