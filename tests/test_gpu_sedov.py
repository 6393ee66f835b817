"""GPU parity for the Sedov blast configuration: 3D shock capturing with the
FV subcell flux, HLLC (a build extension with no reference counterpart) and
the positivity floor (also a build extension).  The oracle carries the same
extensions, so these compare like with like; the LLF-FV variant ties the 3D
shock-capturing path to the reference's own algorithm.

Bars: one RHS relative L2 <= 1e-12 with the identical troubled-element set;
a run of SSPRK3 steps relative L2 <= 1e-10.  At full size (32^3, N=3) the
checks are size-independent: conserved totals and positivity."""
import numpy as np
import pytest

import dgsem_amd as dg
from gpu_pairs import gpu_rhs, make_pair, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
STEPS_TOL = 1e-10
BOX = ([-1.0] * 3, [1.0] * 3)


def sedov_dt(u, N, dx, cfl=0.3, gamma=1.4):
    """dt = CFL dx / (3 (2N+1) lambda_max), lambda_max the largest |v|+c."""
    rho = u[..., 0]
    v = u[..., 1:4] / rho[..., None]
    p = (gamma - 1) * (u[..., 4] - 0.5 * rho * np.sum(v * v, axis=-1))
    lam = np.max(np.linalg.norm(v, axis=-1) + np.sqrt(gamma * p / rho))
    return cfl * dx / (3 * (2 * N + 1) * lam)


def pressure(u, gamma=1.4):
    return (gamma - 1) * (u[..., 4] - 0.5 * np.sum(u[..., 1:4] ** 2, axis=-1) / u[..., 0])


@pytest.mark.parametrize("geometry", [dg.GEOMETRY_NODAL, dg.GEOMETRY_AFFINE])
@pytest.mark.parametrize("fv,surface", [(dg.FLUX_HLLC, dg.FLUX_LLF), (dg.FLUX_LLF, dg.FLUX_LLF),
                                        (dg.FLUX_HLLC, dg.FLUX_HLLC)])
@pytest.mark.parametrize("N,cells", [(3, (6, 6, 6)), (2, (5, 6, 4)), (4, (4, 4, 4))])
def test_sedov_rhs_and_troubled_set(gpu, N, cells, fv, surface, geometry):
    print(f"case sedov rhs N={N} cells={cells} fv={fv} surface={surface} geometry={geometry}",
          flush=True)
    oc, gs = make_pair(3, N, cells, *BOX, geometry=geometry, volume=dg.SHOCK_CAPTURING,
                       surface=surface, fv=fv, amax=0.5, amin=0.001, floor=1e-13)
    u = oc.init_sedov()
    ref, alpha_ref = oc.rhs(u, with_alpha=True)
    got = gpu_rhs(gs, u)
    alpha = gs.alpha().cpu().numpy()
    assert (alpha_ref > 0).any() and not (alpha_ref > 0).all()
    assert np.array_equal(alpha > 0, alpha_ref > 0)
    assert np.max(np.abs(alpha - alpha_ref)) <= 1e-14
    assert rel_l2(got, ref) <= RHS_TOL


@pytest.mark.parametrize("floor", [1e-13, 1e-5])
def test_sedov_ssprk3_steps(gpu, floor):
    """30 steps of the blast.  With floor 1e-5 (the background pressure) the
    clamp acts at the nodes whose pressure dips below it, so the floor's
    energy correction is exercised as well."""
    N, cells, steps = 3, (8, 8, 8), 30
    print(f"case sedov steps floor={floor}", flush=True)
    oc, gs = make_pair(3, N, cells, *BOX, volume=dg.SHOCK_CAPTURING, fv=dg.FLUX_HLLC, amax=0.5,
                       amin=0.001, floor=floor)
    u0 = oc.init_sedov()
    dt = 3 * sedov_dt(u0, N, 2.0 / cells[0])
    ref = oc.ssprk3(u0, dt, steps)
    if floor > 1e-6:
        unclamped = O.Case(3, N, cells, *BOX, volume=O.SHOCK_CAPTURING, fv_flux=O.FLUX_HLLC,
                           alpha_max=0.5, alpha_min=0.001).ssprk3(u0, dt, steps)
        assert pressure(unclamped).min() < floor  # the clamp has work to do
        assert pressure(ref).min() >= floor * (1 - 1e-12)
    u = gs.from_reference_layout(u0)
    t1, t2 = gs.empty(), gs.empty()
    gs.ssprk3_steps(u, dt, steps, t1, t2)
    got = gs.to_reference_layout(u)
    assert rel_l2(got, ref) <= STEPS_TOL
    assert pressure(got).min() >= min(floor, pressure(ref).min()) * (1 - 1e-12)


@pytest.mark.parametrize("geometry", [dg.GEOMETRY_NODAL, dg.GEOMETRY_AFFINE], ids=["nodal", "affine"])
def test_sedov_full_size_conservation(gpu, geometry):
    """32^3, N=3, 100 steps on the GPU alone, with the per-node metric terms
    (read from the box table, as in the benchmark) and with the affine ones.
    With the floor inactive the scheme conserves mass, momentum and energy:
    the quadrature totals sum(w_i w_j w_k / |1/J| u) stay at their initial
    values to round-off.  (The pair terms of each line telescope by the SBP
    property with any symmetric normal, and each interface flux is evaluated
    with one normal for both neighbours.)"""
    N, cells, steps = 3, (32, 32, 32), 100
    print(f"case sedov full size geometry={geometry}", flush=True)
    cfg = dg.SolverConfig(dim=3, polydeg=N, cells=cells, domain_left=BOX[0],
                          domain_right=BOX[1], volume_integral=dg.SHOCK_CAPTURING,
                          fv_flux=dg.FLUX_HLLC, alpha_max=0.5, alpha_min=0.001,
                          positivity_floor=1e-13, geometry=geometry)
    gs = dg.StructuredSolver(cfg)
    u0 = O.Case(3, N, cells, *BOX).init_sedov()  # the initial state's bits only
    w = O.basis(N)["weights"]
    wq = np.einsum("k,j,i->kji", w, w, w).ravel()
    vol = wq[None, :] / np.abs(gs.inverse_jacobian())

    def totals(u):
        return np.einsum("ed,edv->v", vol, u)

    dt = sedov_dt(u0, N, 2.0 / cells[0])
    u = gs.from_reference_layout(u0)
    t1, t2 = gs.empty(), gs.empty()
    gs.ssprk3_steps(u, dt, steps, t1, t2)
    got = gs.to_reference_layout(u)
    assert np.isfinite(got).all()
    assert got[..., 0].min() > 0 and pressure(got).min() > 0
    t0, t = totals(u0), totals(got)
    scale = np.abs(t0).max()
    assert np.max(np.abs(t - t0)) <= 1e-12 * scale
    assert np.all(gs.alpha().cpu().numpy() >= 0)
