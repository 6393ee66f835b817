"""Box meshes read their nodal metric terms from the compact table (the
line kernels' ja_box): the same values as the per-node arrays, so the RHS
and stages must be bit for bit those of the same mesh run from the per-node
arrays.  dgsem_set_node_coordinates with the box's own node coordinates
rebuilds the per-node arrays (bit-identical to the host construction, as
tests/test_gpu_curvilinear.py shows) and turns the table off, which gives
the second run."""
import math

import numpy as np
import pytest
import torch

import dgsem_amd as dg
from dgsem_amd import ensemble, sedov, tgv

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi


def pair_run(cfg, state_fn, stage_dt):
    s = dg.StructuredSolver(cfg)
    u = s.from_reference_layout(state_fn(s.node_coordinates()))
    du_box, st_box = s.empty(), s.empty()
    s.calc_rhs(u, du_box)
    s.ssprk3_stage(1, stage_dt, None, u, st_box)
    coords = torch.from_numpy(s.node_coordinates()).cuda()
    s.set_node_coordinates(coords)
    du, st = s.empty(), s.empty()
    s.calc_rhs(u, du)
    s.ssprk3_stage(1, stage_dt, None, u, st)
    torch.cuda.synchronize()
    return du_box, du, st_box, st


CASES = [
    (3, 3, (4, 4, 4), dg.SPLIT_FORM, "tgv"),
    (3, 3, (5, 3, 4), dg.SPLIT_FORM, "tgv"),
    (3, 2, (3, 4, 2), dg.SPLIT_FORM, "tgv"),
    (3, 4, (3, 2, 3), dg.SPLIT_FORM, "tgv"),
    (3, 5, (2, 3, 2), dg.SPLIT_FORM, "tgv"),
    (3, 3, (6, 6, 6), dg.SHOCK_CAPTURING, "sedov"),
    (2, 3, (8, 6), dg.SHOCK_CAPTURING, "khi"),
    (2, 4, (5, 7), dg.SPLIT_FORM, "khi"),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}d-N{c[1]}-{c[4]}-{c[2]}" for c in CASES])
def test_box_table_is_bitwise_the_nodal_arrays(gpu, case):
    dim, N, cells, vol, kind = case
    print(f"case box table {case}", flush=True)
    if kind == "tgv":
        box, fn = ([0.0] * 3, [TWO_PI] * 3), tgv.tgv_state
    elif kind == "sedov":
        box, fn = ([-1.0] * 3, [1.0] * 3), sedov.sedov_state
    else:
        box, fn = ([-1.0] * 2, [1.0] * 2), ensemble.khi_state
    cfg = dg.SolverConfig(dim=dim, polydeg=N, cells=cells, domain_left=box[0], domain_right=box[1],
                          volume_integral=vol, fv_flux=dg.FLUX_HLLC if kind == "sedov" else dg.FLUX_LLF,
                          alpha_max=0.5 if kind == "sedov" else 0.002, alpha_min=0.001)
    du_box, du, st_box, st = pair_run(cfg, fn, 1e-4)
    assert torch.equal(du_box, du)
    assert torch.equal(st_box, st)


def test_box_table_full_size_tgv(gpu):
    """TGV 64^3 elements, N = 3: the benchmark case."""
    print("case box table full size 64^3 N=3", flush=True)
    cfg = dg.SolverConfig(dim=3, polydeg=3, cells=(64, 64, 64), domain_left=(0, 0, 0),
                          domain_right=(TWO_PI,) * 3)
    du_box, du, st_box, st = pair_run(cfg, tgv.tgv_state, 2e-4)
    assert torch.equal(du_box, du)
    assert torch.equal(st_box, st)
