"""GPU parity at the margins of the parity bars, each case printing the
relative L2 it measured:

  * the fast line kernel (N <= 5: hardware reciprocal with Newton
    refinement, reassociated flux arithmetic) on the hardest states --
    the anisotropic 3x2x4 Mach-0.1 Taylor-Green vortex, whose RHS is some
    700 times smaller than the pressure terms it cancels, and the rough
    state whose logarithmic means leave their series branch -- at every
    degree it serves;
  * the Kelvin-Helmholtz configuration (32^2, N = 3, alpha_max 0.002,
    alpha_min 0.001) at its own CFL-0.5 time step for 100 SSPRK3 steps, with
    the troubled-element set compared with the oracle's after every step;
  * the positivity floor with its density and pressure branches both
    acting (still gas at densities and pressures below the floor), in the
    line kernel and the node kernel;
  * 3D N = 7 shock capturing (the line kernel at its highest degree).

Bars: one RHS relative L2 <= 1e-12 with identical troubled sets; 100 steps
<= 1e-10."""
import math

import numpy as np
import pytest

import dgsem_amd as dg
from dgsem_amd import ensemble
from gpu_pairs import gpu_rhs, make_pair, rel_l2
from test_gpu_high_order import rough_state

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi
RHS_TOL = 1e-12
STEPS_TOL = 1e-10
FLOOR = 1e-13


@pytest.mark.parametrize("geometry", [dg.GEOMETRY_NODAL, dg.GEOMETRY_AFFINE])
@pytest.mark.parametrize("surface", [dg.FLUX_LLF, dg.FLUX_HLLC])
@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_line_kernel_anisotropic_tgv(gpu, N, surface, geometry):
    oc, gs = make_pair(3, N, (3, 2, 4), [0, 0, 0], [TWO_PI] * 3, geometry=geometry, surface=surface)
    u = oc.init_tgv()
    err = rel_l2(gpu_rhs(gs, u), oc.rhs(u))
    print(f"case anisotropic tgv N={N} surface={surface} geometry={geometry}: rel L2 {err:.3e}",
          flush=True)
    assert err <= RHS_TOL


@pytest.mark.parametrize("N", [2, 4, 5])
def test_line_kernel_rough_state(gpu, N):
    oc, gs = make_pair(3, N, (3, 2, 2), [0, 0, 0], [1, 1, 1])
    u = rough_state(oc, 7 + N)
    err = rel_l2(gpu_rhs(gs, u), oc.rhs(u))
    print(f"case rough N={N}: rel L2 {err:.3e}", flush=True)
    assert err <= RHS_TOL


def test_khi_cfl_dt_100_steps_troubled_sets(gpu):
    """C0 at dt = 0.5 dx / ((2N+1) lambda_max) from the initial state; each
    path advances its own state and after every step the troubled set of
    the RHS at that state is compared."""
    N, cells = 3, (32, 32)
    oc, gs = make_pair(2, N, cells, [-1, -1], [1, 1], volume=dg.SHOCK_CAPTURING, amax=0.002,
                       amin=0.001)
    u0 = oc.init_khi()
    dt = ensemble.khi_dt(u0, N, cells)
    assert 1.5e-3 < dt < 3e-3
    print(f"case khi 32^2 N=3 dt={dt:.6e}", flush=True)
    ref = u0.copy()
    u = gs.from_reference_layout(u0)
    t1, t2, du = gs.empty(), gs.empty(), gs.empty()
    worst = 0.0
    for step in range(100):
        ref = oc.ssprk3(ref, dt, 1)
        gs.ssprk3_steps(u, dt, 1, t1, t2)
        _, alpha_ref = oc.rhs(ref, with_alpha=True)
        gs.calc_rhs(u, du)
        alpha = gs.alpha().cpu().numpy()
        assert np.array_equal(alpha > 0, alpha_ref > 0), f"troubled sets differ after step {step + 1}"
        worst = max(worst, rel_l2(gs.to_reference_layout(u), ref))
    err = rel_l2(gs.to_reference_layout(u), ref)
    print(f"  troubled elements at the end {int((alpha_ref > 0).sum())}, "
          f"rel L2 after 100 steps {err:.3e} (worst step {worst:.3e})", flush=True)
    assert (alpha_ref > 0).any()
    assert err <= STEPS_TOL


def slab_state(oc, gamma=1.4):
    """Still gas in three slabs along x: rho = 1, p = 1; rho = 5e-14, p = 1;
    rho = 5e-14, p = 5e-14 (slab by element centre)."""
    x = oc.coordinates()[..., 0]
    xc = np.repeat(x.mean(axis=1, keepdims=True), x.shape[1], axis=1)
    lo, hi = x.min(), x.max()
    third = (hi - lo) / 3.0
    rho = np.where(xc < lo + third, 1.0, 5e-14)
    p = np.where(xc > hi - third, 5e-14, 1.0)
    u = np.zeros(oc.shape)
    u[..., 0] = rho
    u[..., -1] = p / (gamma - 1.0)
    return u


def pressure(u, gamma=1.4):
    d = u.shape[-1] - 2
    return (gamma - 1) * (u[..., -1] - 0.5 * np.sum(u[..., 1:1 + d] ** 2, axis=-1) / u[..., 0])


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 3), (3, 4), (2, 7), (3, 7)])
def test_floor_branches_match_oracle(gpu, dim, N):
    """The stage kernels' floor against the oracle where both of its
    branches act: after every one of 10 steps the clamped-density and
    raised-pressure node sets are identical and the states agree."""
    cells = (6, 2) if dim == 2 else (6, 2, 2)
    box = ([0.0] * dim, [TWO_PI] * dim)
    oc, gs = make_pair(dim, N, cells, *box, floor=FLOOR)
    u0 = slab_state(oc)
    dt = 1e-9
    ref = u0.copy()
    u = gs.from_reference_layout(u0)
    t1, t2 = gs.empty(), gs.empty()
    for step in range(10):
        ref = oc.ssprk3(ref, dt, 1)
        gs.ssprk3_steps(u, dt, 1, t1, t2)
        got = gs.to_reference_layout(u)
        assert np.array_equal(got[..., 0] == FLOOR, ref[..., 0] == FLOOR), step
        assert np.array_equal(pressure(got) <= 1.01 * FLOOR, pressure(ref) <= 1.01 * FLOOR), step
    err = rel_l2(got, ref)
    n_rho = int((ref[..., 0] == FLOOR).sum())
    n_p = int((pressure(ref) <= 1.01 * FLOOR).sum())
    print(f"case floor dim={dim} N={N}: {n_rho} nodes at the density floor, {n_p} at the "
          f"pressure floor, rel L2 after 10 steps {err:.3e}", flush=True)
    assert n_rho > 0 and n_p > 0
    assert got[..., 0].min() >= FLOOR
    assert err <= STEPS_TOL


@pytest.mark.parametrize("geometry", [dg.GEOMETRY_NODAL, dg.GEOMETRY_AFFINE])
@pytest.mark.parametrize("fv", [dg.FLUX_LLF, dg.FLUX_HLLC])
def test_sedov_n7_shock_capturing(gpu, fv, geometry):
    """3D N = 7 with shock capturing runs the line kernel (the node kernel
    serves split form only)."""
    oc, gs = make_pair(3, 7, (2, 2, 2), [-1] * 3, [1] * 3, geometry=geometry,
                       volume=dg.SHOCK_CAPTURING, fv=fv, amax=0.5, amin=0.001, floor=1e-13)
    u = oc.init_sedov()
    ref, alpha_ref = oc.rhs(u, with_alpha=True)
    got = gpu_rhs(gs, u)
    alpha = gs.alpha().cpu().numpy()
    err = rel_l2(got, ref)
    print(f"case sedov N=7 fv={fv} geometry={geometry}: rel L2 {err:.3e}, "
          f"troubled {int((alpha_ref > 0).sum())}/8", flush=True)
    assert (alpha_ref > 0).any()
    assert np.array_equal(alpha > 0, alpha_ref > 0)
    assert np.max(np.abs(alpha - alpha_ref)) <= 1e-14
    assert err <= RHS_TOL
