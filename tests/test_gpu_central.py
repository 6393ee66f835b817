"""GPU parity of CentralFlux (numerical_flux.hpp:230-258) as the split-form
volume flux, the surface flux and the subcell flux of shock capturing, on
2D meshes (the reference defines its normal form for 2D only), against the
oracle, whose central flux and its wiring are pinned in
test_oracle_central.py.  The configurations run the line kernels'
instantiation that selects the fluxes at run time.

Bars: one RHS relative L2 <= 1e-12 (with the identical troubled-element set
under shock capturing); 20 SSPRK3 steps relative L2 <= 1e-11.  The
configurations the library does not build (3D, N = 7, batches, derivatives)
are refused with DGSEM_E_UNSUPPORTED."""
import numpy as np
import pytest
import torch

import dgsem_amd as dg
from gpu_pairs import gpu_rhs, make_pair, rel_l2
from oracle import oracle as O
from test_gpu_boundaries import INFLOW, state

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
STEPS_TOL = 1e-11
C, CH = dg.FLUX_CENTRAL, dg.FLUX_CHANDRASHEKAR

# (volume flux, surface flux, volume integral, subcell flux)
FLUXES = {
    "central-central": (C, C, dg.SPLIT_FORM, dg.FLUX_LLF),
    "central-llf": (C, dg.FLUX_LLF, dg.SPLIT_FORM, dg.FLUX_LLF),
    "central-hllc": (C, dg.FLUX_HLLC, dg.SPLIT_FORM, dg.FLUX_LLF),
    "chandrashekar-central": (CH, C, dg.SPLIT_FORM, dg.FLUX_LLF),
    "sc-central-volume": (C, dg.FLUX_HLLC, dg.SHOCK_CAPTURING, dg.FLUX_LLF),
    "sc-central-subcells": (CH, dg.FLUX_LLF, dg.SHOCK_CAPTURING, C),
    "sc-all-central": (C, C, dg.SHOCK_CAPTURING, C),
}


def khi_like(x, gamma=1.4):
    """A Kelvin-Helmholtz-like shear layer with a smooth perturbation."""
    X, Y = x[..., 0], x[..., 1]
    B = np.tanh(15 * Y + 7.5) - np.tanh(15 * Y - 7.5)
    rho = 0.5 + 0.75 * B
    vx = 0.5 * (B - 1) + 0.02 * np.cos(np.pi * X + 0.3)
    vy = 0.1 * np.sin(2 * np.pi * X)
    p = 1.0 + 0.05 * np.cos(2 * np.pi * X + 0.4) * np.sin(2 * np.pi * Y + 0.1)
    return np.stack([rho, rho * vx, rho * vy, p / (gamma - 1) + 0.5 * rho * (vx * vx + vy * vy)], -1)


@pytest.mark.parametrize("fluxes", list(FLUXES))
@pytest.mark.parametrize("N,cells", [(2, (6, 5)), (3, (5, 6)), (4, (4, 5)), (5, (4, 3))])
def test_central_rhs_periodic(gpu, N, cells, fluxes):
    vflux, surface, volume, fv = FLUXES[fluxes]
    print(f"case central rhs {fluxes} N={N} cells={cells}", flush=True)
    for geometry in (dg.GEOMETRY_NODAL, dg.GEOMETRY_AFFINE):
        oc, gs = make_pair(2, N, cells, [-1.0, -1.0], [1.0, 1.0], geometry=geometry, volume=volume,
                           surface=surface, fv=fv, amax=0.5, amin=0.001, vflux=vflux)
        u = khi_like(oc.coordinates())
        ref, alpha = oc.rhs(u, with_alpha=True)
        got = gpu_rhs(gs, u)
        err = rel_l2(got, ref)
        print(f"  {geometry}: rel {err:.2e}", flush=True)
        assert err <= RHS_TOL, geometry
        if volume == dg.SHOCK_CAPTURING:
            a = gs.alpha().cpu().numpy()
            assert np.array_equal(a > 0, alpha > 0)
            assert 0 < int((alpha > 0).sum()) < oc.E


@pytest.mark.parametrize("fluxes", ["central-central", "sc-all-central"])
def test_central_ssprk3_steps(gpu, fluxes):
    vflux, surface, volume, fv = FLUXES[fluxes]
    N, cells, dt, steps = 3, (8, 8), 2e-4, 20
    print(f"case central steps {fluxes}", flush=True)
    oc, gs = make_pair(2, N, cells, [-1.0, -1.0], [1.0, 1.0], volume=volume, surface=surface, fv=fv,
                       amax=0.5, amin=0.001, vflux=vflux)
    u0 = khi_like(oc.coordinates())
    ref = oc.ssprk3(u0, dt, steps)
    u = gs.from_reference_layout(u0)
    t1, t2 = gs.empty(), gs.empty()
    gs.ssprk3_steps(u, dt, steps, t1, t2)
    torch.cuda.synchronize()
    err = rel_l2(gs.to_reference_layout(u), ref)
    print(f"  rel {err:.2e}", flush=True)
    assert err <= STEPS_TOL


@pytest.mark.parametrize("N,cells", [(2, (6, 5)), (4, (5, 4))])
def test_central_surface_flux_with_boundaries(gpu, N, cells):
    """Slip walls, Dirichlet inflow and outflow faces take the central flux
    with their outer states, as the periodic faces do."""
    faces = [dg.BC_DIRICHLET, dg.BC_OUTFLOW, dg.BC_SLIP_WALL, dg.BC_SLIP_WALL]
    left, right = (-1.0, -0.5), (1.0, 0.75)
    print(f"case central boundaries N={N}", flush=True)
    for geometry in (dg.GEOMETRY_NODAL, dg.GEOMETRY_AFFINE):
        oc = O.Case(2, N, cells, list(left), list(right), periodic=[0, 0], surface_flux=O.FLUX_CENTRAL,
                    volume_flux=O.VOLUME_CENTRAL)
        states = [INFLOW if k == dg.BC_DIRICHLET else None for k in faces]
        for f, kind in enumerate(faces):
            oc.set_boundary(f, kind, state=states[f])
        cfg = dg.SolverConfig(dim=2, polydeg=N, cells=cells, domain_left=list(left),
                              domain_right=list(right), volume_flux=C, surface_flux=C, geometry=geometry,
                              face_boundary=faces, face_state=[s or [0.0] * 4 for s in states])
        gs = dg.StructuredSolver(cfg)
        if geometry == dg.GEOMETRY_AFFINE:
            h = [(right[k] - left[k]) / cells[k] / 2.0 for k in range(2)]
            Ja = np.zeros((oc.E, oc.ndofs, 2, 2))
            Ja[..., 0, 0], Ja[..., 1, 1] = h[1], h[0]
            oc.set_metrics(Ja, np.full((oc.E, oc.ndofs), 1.0 / (h[0] * h[1])))
        u = state(oc, seed=N, jump=False)
        err = rel_l2(gpu_rhs(gs, u), oc.rhs(u))
        print(f"  {geometry}: rel {err:.2e}", flush=True)
        assert err <= RHS_TOL


def test_central_navier_stokes(gpu):
    """The hyperbolic part of Navier-Stokes with the central fluxes."""
    print("case central navier-stokes", flush=True)
    oc, gs = make_pair(2, 3, (5, 4), [-1.0, -1.0], [1.0, 1.0], surface=C, vflux=C, mu=5e-3)
    u = khi_like(oc.coordinates())
    err = rel_l2(gpu_rhs(gs, u), oc.rhs(u))
    print(f"  rel {err:.2e}", flush=True)
    assert err <= RHS_TOL


def _config(**kw):
    base = dict(dim=2, polydeg=3, cells=(4, 4), domain_left=(-1, -1), domain_right=(1, 1), surface_flux=C)
    base.update(kw)
    return dg.SolverConfig(**base)


@pytest.mark.parametrize("what", ["3d", "n7", "3d-volume"])
def test_central_refused_where_not_built(gpu, what):
    if what == "3d":
        cfg = _config(dim=3, cells=(2, 2, 2), domain_left=(0, 0, 0), domain_right=(1, 1, 1))
    elif what == "3d-volume":
        cfg = _config(dim=3, cells=(2, 2, 2), domain_left=(0, 0, 0), domain_right=(1, 1, 1),
                      surface_flux=dg.FLUX_LLF, volume_flux=C)
    else:
        cfg = _config(polydeg=7)
    with pytest.raises(dg.DgsemError, match="CentralFlux"):
        dg.StructuredSolver(cfg)


def test_central_refused_by_batches_and_derivatives(gpu):
    with pytest.raises(dg.DgsemError, match="LLF or HLLC"):
        dg.Batch([_config(), _config()])
    gs = dg.StructuredSolver(_config())
    u = gs.from_reference_layout(khi_like(gs.node_coordinates()))
    with pytest.raises(dg.DgsemError, match="CentralFlux is not differentiated"):
        gs.rhs_jvp(u, torch.ones_like(u))
