"""GPU parity: the HIP right-hand side and SSPRK3 stages against the CPU
restatement of the reference (oracle/), on the same inputs.

Bars (tolerances written here, from the build's north star):
  * one RHS evaluation: relative L2 <= 1e-12;
  * shock capturing: identical set of troubled elements (alpha > 0);
  * state after 100 SSPRK3 steps: relative L2 <= 1e-10.
NODAL geometry uses the reference's own per-node metric terms; AFFINE uses
exact per-element metrics, so its oracle run is fed the same exact metrics.
"""
import math

import numpy as np
import pytest

import dgsem_amd as dg
from gpu_pairs import gpu_rhs, make_pair, rel_l2

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
STEPS_TOL = 1e-10
TWO_PI = 2.0 * math.pi


@pytest.mark.parametrize("geometry", [dg.GEOMETRY_NODAL, dg.GEOMETRY_AFFINE])
@pytest.mark.parametrize("N,cells", [(3, (4, 4, 4)), (3, (6, 5, 4)), (2, (4, 4, 4)),
                                     (4, (3, 3, 3)), (5, (2, 3, 2)), (7, (2, 2, 2))])
def test_tgv_rhs_matches_oracle(gpu, N, cells, geometry):
    print(f"case tgv N={N} cells={cells} geometry={geometry}", flush=True)
    oc, gs = make_pair(3, N, cells, [0, 0, 0], [TWO_PI] * 3, geometry=geometry)
    u = oc.init_tgv()
    ref = oc.rhs(u)
    got = gpu_rhs(gs, u)
    assert rel_l2(got, ref) <= RHS_TOL


@pytest.mark.parametrize("N,cells", [(3, (8, 8)), (3, (32, 32)), (2, (6, 10)), (4, (8, 8)),
                                     (5, (5, 7)), (7, (4, 4))])
def test_khi_shock_capturing_rhs_and_troubled_set(gpu, N, cells):
    print(f"case khi N={N} cells={cells}", flush=True)
    oc, gs = make_pair(2, N, cells, [-1, -1], [1, 1], volume=dg.SHOCK_CAPTURING,
                       amax=0.002, amin=0.001)
    u = oc.init_khi()
    ref, alpha_ref = oc.rhs(u, with_alpha=True)
    got = gpu_rhs(gs, u)
    alpha = gs.alpha().cpu().numpy()
    assert rel_l2(got, ref) <= RHS_TOL
    assert np.array_equal(alpha > 0, alpha_ref > 0)
    assert np.max(np.abs(alpha - alpha_ref)) <= 1e-14


def test_khi_ssprk3_100_steps(gpu):
    """config-0 style run: 32x32, N=3, alpha_max=0.002, 100 steps."""
    oc, gs = make_pair(2, 3, (32, 32), [-1, -1], [1, 1], volume=dg.SHOCK_CAPTURING,
                       amax=0.002, amin=0.001)
    u0 = oc.init_khi()
    dt = 2e-4
    ref = oc.ssprk3(u0, dt, 100)
    u = gs.from_reference_layout(u0)
    rk = dg.SSPRK3(gs)
    for _ in range(100):
        rk.step(u, dt)
    got = gs.to_reference_layout(u)
    assert rel_l2(got, ref) <= STEPS_TOL


@pytest.mark.parametrize("geometry", [dg.GEOMETRY_NODAL, dg.GEOMETRY_AFFINE])
def test_tgv_ssprk3_100_steps(gpu, geometry):
    oc, gs = make_pair(3, 3, (4, 4, 4), [0, 0, 0], [TWO_PI] * 3, geometry=geometry)
    u0 = oc.init_tgv()
    dt = 1e-3
    ref = oc.ssprk3(u0, dt, 100)
    u = gs.from_reference_layout(u0)
    t1, t2 = gs.empty(), gs.empty()
    gs.ssprk3_steps(u, dt, 100, t1, t2)
    got = gs.to_reference_layout(u)
    assert rel_l2(got, ref) <= STEPS_TOL
