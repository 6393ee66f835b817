"""GPU parity of the non-periodic 2D boundary conditions (slip wall,
outflow, Dirichlet with a constant or a per-node outer state) against the
oracle, whose boundary fluxes are pinned by the properties in
test_oracle_boundaries.py.

Bars: one RHS relative L2 <= 1e-12 with the identical troubled-element set;
100 SSPRK3 steps relative L2 <= 1e-10."""
import numpy as np
import pytest
import torch

import dgsem_amd as dg
from gpu_pairs import gpu_rhs, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
STEPS_TOL = 1e-10
GAMMA = 1.4
INFLOW = [1.4, 1.4 * 0.6, 0.0, 1.0 / (GAMMA - 1) + 0.5 * 1.4 * 0.36]

# face kinds, -x, +x, -y, +y
LAYOUTS = {
    "walls": [dg.BC_SLIP_WALL] * 4,
    "channel": [dg.BC_DIRICHLET, dg.BC_OUTFLOW, dg.BC_SLIP_WALL, dg.BC_SLIP_WALL],
    "x_walls": [dg.BC_SLIP_WALL, dg.BC_SLIP_WALL, dg.BC_PERIODIC, dg.BC_PERIODIC],
    "y_open": [dg.BC_PERIODIC, dg.BC_PERIODIC, dg.BC_DIRICHLET, dg.BC_OUTFLOW],
}


def make_bc_pair(N, cells, faces, *, geometry=dg.GEOMETRY_NODAL, volume=dg.SPLIT_FORM,
                 surface=dg.FLUX_LLF, left=(-1.0, -0.5), right=(1.0, 0.75)):
    periodic = [int(faces[0] == dg.BC_PERIODIC), int(faces[2] == dg.BC_PERIODIC)]
    oc = O.Case(2, N, cells, list(left), list(right), periodic=periodic, volume=volume,
                surface_flux=surface, fv_flux=O.FLUX_LLF, alpha_max=0.5, alpha_min=0.001)
    states = [INFLOW if k == dg.BC_DIRICHLET else None for k in faces]
    for f, kind in enumerate(faces):
        if kind != dg.BC_PERIODIC:
            oc.set_boundary(f, kind, state=states[f])
    cfg = dg.SolverConfig(dim=2, polydeg=N, cells=cells, domain_left=list(left),
                          domain_right=list(right), volume_integral=volume, surface_flux=surface,
                          fv_flux=dg.FLUX_LLF, alpha_max=0.5, alpha_min=0.001, geometry=geometry,
                          face_boundary=faces, face_state=[s or [0.0] * 4 for s in states])
    gs = dg.StructuredSolver(cfg)
    if geometry == dg.GEOMETRY_AFFINE:
        h = [(right[k] - left[k]) / cells[k] / 2.0 for k in range(2)]
        Ja = np.zeros((oc.E, oc.ndofs, 2, 2))
        Ja[..., 0, 0], Ja[..., 1, 1] = h[1], h[0]
        oc.set_metrics(Ja, np.full((oc.E, oc.ndofs), 1.0 / (h[0] * h[1])))
    return oc, gs


def state(oc, seed=0, jump=True):
    """A smooth field with a density/pressure jump across x = 0.1 (so that the
    indicator marks some elements and not others)."""
    rng = np.random.default_rng(seed)
    x = oc.coordinates()
    rho = 1.0 + 0.2 * np.sin(3 * x[..., 0]) * np.cos(2 * x[..., 1]) + 0.01 * rng.random(x.shape[:-1])
    p = 1.0 + 0.1 * np.cos(x[..., 0] + x[..., 1])
    if jump:
        rho = np.where(x[..., 0] < 0.1, rho + 0.8, rho)
        p = np.where(x[..., 0] < 0.1, p + 1.5, p)
    vx = 0.3 * np.cos(2 * x[..., 1]) + 0.05 * rng.random(rho.shape)
    vy = 0.2 * np.sin(3 * x[..., 0])
    return np.stack([rho, rho * vx, rho * vy, p / (GAMMA - 1) + 0.5 * rho * (vx * vx + vy * vy)], -1)


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("N,cells", [(1, (7, 6)), (2, (6, 5)), (3, (6, 5)), (4, (5, 4)),
                                     (5, (4, 4)), (7, (3, 4))])
def test_boundary_rhs_split_form(gpu, N, cells, layout):
    print(f"case bc rhs split N={N} cells={cells} layout={layout}", flush=True)
    for geometry in (dg.GEOMETRY_NODAL, dg.GEOMETRY_AFFINE):
        oc, gs = make_bc_pair(N, cells, LAYOUTS[layout], geometry=geometry)
        u = state(oc, seed=N, jump=False)
        ref = oc.rhs(u)
        got = gpu_rhs(gs, u)
        assert rel_l2(got, ref) <= RHS_TOL, geometry


@pytest.mark.parametrize("surface", [dg.FLUX_LLF, dg.FLUX_HLLC])
@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("N,cells", [(2, (8, 6)), (3, (8, 6)), (5, (6, 4))])
def test_boundary_rhs_shock_capturing(gpu, N, cells, layout, surface):
    print(f"case bc rhs sc N={N} cells={cells} layout={layout} surface={surface}", flush=True)
    oc, gs = make_bc_pair(N, cells, LAYOUTS[layout], volume=dg.SHOCK_CAPTURING, surface=surface)
    u = state(oc, seed=10 + N)
    ref, alpha_ref = oc.rhs(u, with_alpha=True)
    got = gpu_rhs(gs, u)
    alpha = gs.alpha().cpu().numpy()
    assert (alpha_ref > 0).any() and not (alpha_ref > 0).all()
    assert np.array_equal(alpha > 0, alpha_ref > 0)
    assert np.max(np.abs(alpha - alpha_ref)) <= 1e-14
    assert rel_l2(got, ref) <= RHS_TOL


def test_per_node_dirichlet_states(gpu):
    """Per-node outer states on the two x faces (set after creation, read by
    every later RHS), and reverting one face to its constant state."""
    N, cells = 3, (6, 5)
    faces = LAYOUTS["channel"]
    oc, gs = make_bc_pair(N, cells, faces, volume=dg.SHOCK_CAPTURING)
    u = state(oc, seed=7)
    rng = np.random.default_rng(1)
    vals = np.tile(np.asarray(INFLOW), (cells[1] * (N + 1), 1))
    vals[:, 0] += 0.1 * rng.random(len(vals))
    vals[:, 3] += 0.2 * rng.random(len(vals))
    oc.set_boundary(0, O.BC_DIRICHLET, state=INFLOW, values=vals)
    dvals = torch.from_numpy(vals).cuda()
    gs.set_boundary_values(0, dvals)
    ref = oc.rhs(u)
    got = gpu_rhs(gs, u)
    assert rel_l2(got, ref) <= RHS_TOL
    oc.set_boundary(0, O.BC_DIRICHLET, state=INFLOW)
    gs.set_boundary_values(0, None)
    assert rel_l2(gpu_rhs(gs, u), oc.rhs(u)) <= RHS_TOL
    with pytest.raises(dg.DgsemError):
        gs.set_boundary_values(1, dvals)  # the outflow face takes no states


@pytest.mark.parametrize("layout", ["channel", "walls"])
def test_boundary_ssprk3_steps(gpu, layout):
    """100 steps of the jump in a channel (inflow, outflow, walls) and in a
    closed box, shock capturing on."""
    N, cells, steps = 3, (10, 6), 100
    print(f"case bc steps layout={layout}", flush=True)
    oc, gs = make_bc_pair(N, cells, LAYOUTS[layout], volume=dg.SHOCK_CAPTURING)
    u0 = state(oc, seed=3)
    dt = 1e-3
    ref = oc.ssprk3(u0, dt, steps)
    u = gs.from_reference_layout(u0)
    t1, t2 = gs.empty(), gs.empty()
    gs.ssprk3_steps(u, dt, steps, t1, t2)
    got = gs.to_reference_layout(u)
    assert np.isfinite(got).all()
    assert rel_l2(got, ref) <= STEPS_TOL
    if layout == "walls":
        # closed box: mass and energy are conserved
        w = O.basis(N)["weights"]
        vol = np.einsum("j,i->ji", w, w).ravel()[None, :] / np.abs(oc.inverse_jacobian())
        for var in (0, 3):
            t0, t = np.sum(vol * u0[..., var]), np.sum(vol * got[..., var])
            assert abs(t - t0) <= 1e-12 * abs(t0), var


def test_boundary_configurations_rejected(gpu):
    """Batches are built for periodic meshes; they refuse non-periodic cases
    instead of reading missing neighbours.  (The Euler tangent and adjoint
    take them: test_gpu_adjoint_bc.py.)"""
    _, gs = make_bc_pair(2, (4, 4), LAYOUTS["walls"])
    cfg = gs.cfg
    with pytest.raises(dg.DgsemError):
        dg.Batch([cfg, cfg])


@pytest.mark.parametrize("volume", [dg.SPLIT_FORM, dg.SHOCK_CAPTURING])
def test_mixed_dirichlet_slip_wall(gpu, volume):
    """The double-Mach-reflection bottom wall: per node either an inflow
    state or the slip-wall reflection (MixedDirichletSlipWallBC), here on the
    lower y face, with a Dirichlet top and x inflow/outflow."""
    print(f"case bc mixed volume={volume}", flush=True)
    N, cells = 3, (8, 5)
    n = N + 1
    faces = [dg.BC_DIRICHLET, dg.BC_OUTFLOW, dg.BC_MIXED_DIRICHLET_SLIP_WALL, dg.BC_DIRICHLET]
    oc, gs = make_bc_pair(N, cells, faces, volume=volume)
    x_nodes = np.linspace(-1.0, 1.0, cells[0] * n)
    vals = np.column_stack([np.tile(INFLOW, (cells[0] * n, 1)), (x_nodes < -0.4).astype(float)])
    oc.set_boundary(2, O.BC_MIXED_DIRICHLET_SLIP_WALL, values=vals)
    u = state(oc, seed=21)
    with pytest.raises(dg.DgsemError):
        gpu_rhs(gs, u)  # the mixed face's per-node states are required
    gs.set_boundary_values(2, torch.from_numpy(vals).cuda())
    ref = oc.rhs(u)
    assert rel_l2(gpu_rhs(gs, u), ref) <= RHS_TOL
