"""Shared helpers of the GPU parity tests: one oracle case and one GPU solver
built from the same configuration."""
import numpy as np

import dgsem_amd as dg
from oracle import oracle as O


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make_pair(dim, N, cells, left, right, *, geometry=dg.GEOMETRY_NODAL, volume=dg.SPLIT_FORM,
              surface=dg.FLUX_LLF, fv=dg.FLUX_LLF, amax=0.5, amin=0.001, floor=0.0, mu=None,
              vflux=dg.FLUX_CHANDRASHEKAR):
    """Oracle case and GPU solver of one configuration; mu: Navier-Stokes
    with that viscosity and Prandtl number 0.72 (else compressible Euler)."""
    ns = {} if mu is None else dict(eq=O.NAVIER_STOKES, mu=mu, prandtl=0.72)
    oc = O.Case(dim, N, cells, left, right, volume=volume, surface_flux=surface, fv_flux=fv,
                alpha_max=amax, alpha_min=amin, positivity_floor=floor, volume_flux=vflux, **ns)
    ns = {} if mu is None else dict(equation=dg.NAVIER_STOKES, mu=mu, prandtl=0.72)
    cfg = dg.SolverConfig(dim=dim, polydeg=N, cells=cells, domain_left=left, domain_right=right,
                          volume_integral=volume, volume_flux=vflux, surface_flux=surface, fv_flux=fv,
                          alpha_max=amax,
                          alpha_min=amin, positivity_floor=floor, geometry=geometry, **ns)
    gs = dg.StructuredSolver(cfg)
    if geometry == dg.GEOMETRY_AFFINE:
        # feed the oracle the exact per-element metrics the GPU uses
        h = [(right[k] - left[k]) / cells[k] / 2.0 for k in range(dim)]
        Ja = np.zeros((oc.E, oc.ndofs, dim, dim))
        for r in range(dim):
            Ja[..., r, r] = np.prod([h[k] for k in range(dim) if k != r])
        invJ = np.full((oc.E, oc.ndofs), 1.0 / np.prod(h))
        oc.set_metrics(Ja, invJ)
    return oc, gs


def gpu_rhs(gs, u_ref):
    u = gs.from_reference_layout(u_ref)
    du = gs.empty()
    gs.calc_rhs(u, du, 0.0)
    return gs.to_reference_layout(du)
