"""Navier-Stokes with the HLLC flux as the surface flux, the subcell flux of
shock capturing, or both: the forward RHS against the oracle, and the tangent
and adjoint of the RHS (alpha frozen under shock capturing) and of one
SSPRK3 step.  HLLC is a build extension in multiple dimensions (the reference
has it for 1D Euler only), so the oracle's HLLC is the comparison; the BR1
viscous terms follow numerical_flux.hpp / parabolic_viscous.hpp as in the
LLF configurations.

The states keep HLLC's contact speed away from zero on every face (a uniform
drift, as in test_gpu_adjoint_sc.py), where the flux switches from the left
star state to the right one and a central difference would straddle the
switch.

Bars: RHS relative L2 <= 1e-12 with the identical troubled set; tangent
against central differences (h = 1e-7) of the oracle's RHS, alpha
overridden to the same frozen values, relative 1e-6; <w, J v> = <J^T w, v>
to 1e-12 relative, for the RHS and for one step."""
import numpy as np
import pytest
import torch

import dgsem_amd as dg
from gpu_pairs import gpu_rhs, make_pair, rel_l2
from test_gpu_adjoint_sc import direction, drifting_bump, perturbed_khi

pytestmark = pytest.mark.gpu

H, L = dg.FLUX_HLLC, dg.FLUX_LLF
MU = 0.02

CASES = [
    # dim, N, cells, volume integral, surface flux, subcell flux, geometry
    (2, 3, (6, 5), dg.SPLIT_FORM, H, L, dg.GEOMETRY_NODAL),
    (2, 3, (8, 8), dg.SHOCK_CAPTURING, L, H, dg.GEOMETRY_NODAL),
    (2, 4, (6, 5), dg.SHOCK_CAPTURING, H, H, dg.GEOMETRY_AFFINE),
    (3, 2, (3, 3, 3), dg.SPLIT_FORM, H, L, dg.GEOMETRY_AFFINE),
    (3, 3, (4, 4, 4), dg.SHOCK_CAPTURING, L, H, dg.GEOMETRY_NODAL),
    (3, 3, (4, 4, 4), dg.SHOCK_CAPTURING, H, H, dg.GEOMETRY_NODAL),
]


def ids(c):
    name = {H: "hllc", L: "llf"}
    vol = "sc" if c[3] == dg.SHOCK_CAPTURING else "split"
    return (f"{c[0]}d-N{c[1]}-{vol}-surface_{name[c[4]]}-subcells_{name[c[5]]}-"
            f"{'nodal' if c[6] == dg.GEOMETRY_NODAL else 'affine'}")


def pair(case):
    dim, N, cells, volume, surface, fv, geo = case
    box = ([-1.0] * dim, [1.0] * dim)
    oc, gs = make_pair(dim, N, cells, *box, geometry=geo, volume=volume, surface=surface, fv=fv,
                       amax=0.5, amin=0.001, mu=MU)
    x = oc.coordinates()
    u = perturbed_khi(x, drift=(0.7, 0.3)) if dim == 2 else drifting_bump(x)
    return oc, gs, u


@pytest.mark.parametrize("case", CASES, ids=[ids(c) for c in CASES])
def test_rhs_against_oracle(gpu, case):
    print(f"case ns-hllc rhs {ids(case)}", flush=True)
    oc, gs, u = pair(case)
    ref, alpha = oc.rhs(u, with_alpha=True)
    err = rel_l2(gpu_rhs(gs, u), ref)
    print(f"  rel {err:.2e}", flush=True)
    assert err <= 1e-12
    if case[3] == dg.SHOCK_CAPTURING:
        a = gs.alpha().cpu().numpy()
        assert np.array_equal(a > 0, alpha > 0)
        assert 0 < int((alpha > 0).sum()) < oc.E


@pytest.mark.parametrize("case", CASES, ids=[ids(c) for c in CASES])
def test_tangent_matches_oracle_finite_differences(gpu, case):
    print(f"case ns-hllc jvp fd {ids(case)}", flush=True)
    oc, gs, u0 = pair(case)
    sc = case[3] == dg.SHOCK_CAPTURING
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


@pytest.mark.parametrize("case", CASES, ids=[ids(c) for c in CASES])
def test_adjoint_dot_product(gpu, case):
    print(f"case ns-hllc vjp dot {ids(case)}", flush=True)
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


@pytest.mark.parametrize("case", [CASES[2], CASES[5]], ids=[ids(CASES[2]), ids(CASES[5])])
def test_step_adjoint_dot_product(gpu, case):
    print(f"case ns-hllc step dot {ids(case)}", flush=True)
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
