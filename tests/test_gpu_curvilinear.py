"""Curvilinear meshes on the GPU: the element geometry computed by the GPU
from mapped node coordinates (dgsem_set_node_coordinates) equals the
oracle's (pinned in test_oracle_curvilinear.py) bit for bit, and the RHS and
SSPRK3 steps on warped meshes match the oracle to the usual bars
(RHS <= 1e-12 relative L2 with the identical troubled set; steps <= 1e-10)."""
import math

import numpy as np
import pytest

import dgsem_amd as dg
from gpu_pairs import gpu_rhs, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def warp(ref, a=0.1):
    s = np.prod(np.sin(np.pi * ref), axis=-1)
    out = ref.copy()
    for k in range(ref.shape[-1]):
        out[..., k] = ref[..., k] + a * s * (1.0 + 0.5 * k)
    return out


def warped_pair(dim, N, cells, **kw):
    eq = kw.pop("eq", dg.EULER)
    volume = kw.pop("volume", dg.SPLIT_FORM)
    amax = kw.pop("amax", 0.5)
    oc = O.Case(dim, N, cells, [-1.0] * dim, [1.0] * dim, eq=eq, volume=volume, alpha_max=amax,
                fv_flux=kw.get("fv", dg.FLUX_LLF))
    cfg = dg.SolverConfig(dim=dim, polydeg=N, cells=cells, domain_left=[-1.0] * dim,
                          domain_right=[1.0] * dim, equation=eq, volume_integral=volume,
                          alpha_max=amax, fv_flux=kw.get("fv", dg.FLUX_LLF))
    gs = dg.StructuredSolver(cfg)
    ref = gs.reference_coordinates()
    assert np.array_equal(ref, oc.reference_coordinates())
    gs.apply_mapping(warp)
    oc.set_coordinates(warp(ref))
    return oc, gs


@pytest.mark.parametrize("dim,N,cells", [(2, 3, (5, 4)), (2, 7, (2, 3)), (3, 2, (3, 2, 2)),
                                         (3, 3, (2, 3, 2)), (3, 7, (2, 1, 2))])
def test_device_geometry_equals_oracle(gpu, dim, N, cells):
    print(f"case curvilinear geometry dim={dim} N={N}", flush=True)
    oc, gs = warped_pair(dim, N, cells, volume=dg.SHOCK_CAPTURING)
    assert np.array_equal(gs.node_coordinates(), oc.coordinates())
    assert np.array_equal(gs.contravariant_vectors(), oc.contravariant())
    assert np.array_equal(gs.inverse_jacobian(), oc.inverse_jacobian())


def test_device_geometry_equals_host_construction(gpu):
    """The linear mapping's coordinates through the GPU path reproduce the
    geometry the library built on the host at creation, bit for bit."""
    for dim, N, cells in ((2, 4, (3, 4)), (3, 3, (2, 2, 3))):
        cfg = dg.SolverConfig(dim=dim, polydeg=N, cells=cells, domain_left=[0.0] * dim,
                              domain_right=[2 * math.pi] * dim, volume_integral=dg.SHOCK_CAPTURING)
        gs = dg.StructuredSolver(cfg)
        x, ja, ij = gs.node_coordinates(), gs.contravariant_vectors(), gs.inverse_jacobian()
        oc = O.Case(dim, N, cells, [0.0] * dim, [2 * math.pi] * dim)
        u = gs.from_reference_layout(oc.init_khi() if dim == 2 else oc.init_tgv())
        r0 = gs.to_reference_layout(_rhs(gs, u))
        assert np.isfinite(r0).all()
        gs.set_node_coordinates(x)
        assert np.array_equal(gs.contravariant_vectors(), ja)
        assert np.array_equal(gs.inverse_jacobian(), ij)
        assert np.array_equal(gs.to_reference_layout(_rhs(gs, u)), r0)


def _rhs(gs, u):
    du = gs.empty()
    gs.calc_rhs(u, du, 0.0)
    return du


@pytest.mark.parametrize("dim,N,cells", [(2, 2, (6, 5)), (2, 3, (5, 4)), (2, 5, (3, 3)),
                                         (2, 7, (2, 3)), (3, 2, (3, 3, 2)), (3, 3, (3, 2, 3)),
                                         (3, 4, (2, 2, 2)), (3, 7, (2, 1, 2))])
def test_curvilinear_rhs(gpu, dim, N, cells):
    print(f"case curvilinear rhs dim={dim} N={N}", flush=True)
    oc, gs = warped_pair(dim, N, cells)
    u = oc.init_khi() if dim == 2 else oc.init_tgv()
    assert rel_l2(gpu_rhs(gs, u), oc.rhs(u)) <= 1e-12
    # free stream stays put on the GPU too
    v = np.zeros(oc.shape)
    v[..., 0] = 1.1
    v[..., 1] = 0.33
    v[..., dim + 1] = 2.5 + 0.5 * 0.33 ** 2 / 1.1
    assert np.max(np.abs(gpu_rhs(gs, v))) <= 1e-11


@pytest.mark.parametrize("dim,N,cells", [(2, 3, (8, 6)), (3, 3, (4, 3, 3))])
def test_curvilinear_shock_capturing(gpu, dim, N, cells):
    print(f"case curvilinear sc dim={dim}", flush=True)
    oc, gs = warped_pair(dim, N, cells, volume=dg.SHOCK_CAPTURING, fv=dg.FLUX_HLLC)
    u = oc.init_khi() if dim == 2 else oc.init_sedov(sigma=0.0625)
    ref, alpha_ref = oc.rhs(u, with_alpha=True)
    got = gpu_rhs(gs, u)
    alpha = gs.alpha().cpu().numpy()
    assert (alpha_ref > 0).any() and not (alpha_ref > 0).all()
    assert np.array_equal(alpha > 0, alpha_ref > 0)
    assert rel_l2(got, ref) <= 1e-12


@pytest.mark.parametrize("dim,N,cells", [(2, 3, (5, 4)), (2, 7, (2, 3)), (3, 7, (2, 1, 2))])
def test_curvilinear_navier_stokes(gpu, dim, N, cells):
    print(f"case curvilinear ns dim={dim} N={N}", flush=True)
    oc, gs = warped_pair(dim, N, cells, eq=dg.NAVIER_STOKES)
    u = oc.init_khi() if dim == 2 else oc.init_tgv()
    assert rel_l2(gpu_rhs(gs, u), oc.rhs(u)) <= 1e-12


def test_curvilinear_steps(gpu):
    print("case curvilinear steps", flush=True)
    oc, gs = warped_pair(2, 3, (10, 8), volume=dg.SHOCK_CAPTURING, amax=0.002)
    u0 = oc.init_khi()
    dt, steps = 5e-4, 60
    ref = oc.ssprk3(u0, dt, steps)
    u = gs.from_reference_layout(u0)
    t1, t2 = gs.empty(), gs.empty()
    gs.ssprk3_steps(u, dt, steps, t1, t2)
    assert rel_l2(gs.to_reference_layout(u), ref) <= 1e-10


def test_affine_mode_refuses_mapped_coordinates(gpu):
    cfg = dg.SolverConfig(dim=2, polydeg=3, cells=(2, 2), domain_left=[-1, -1], domain_right=[1, 1],
                          geometry=dg.GEOMETRY_AFFINE)
    gs = dg.StructuredSolver(cfg)
    with pytest.raises(dg.DgsemError):
        gs.apply_mapping(warp)
