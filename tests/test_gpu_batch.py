"""Batched ensembles: every case advanced by the persistent batch kernel
(one launch per SSPRK3 stage, replayed from a HIP graph) must be
bit-identical to the same case advanced alone by its single-case handle, for
any tuning the autotuner may pick.  A batch RHS is also checked against the
oracle for a few cases, which ties the batch to the reference's algorithm."""
import numpy as np
import pytest
import torch

import dgsem_amd as dg
from dgsem_amd import ensemble
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def single_runs(batch, u0, dts, steps):
    out = []
    for i in range(batch.ncases):
        s = batch.case(i)
        u = batch.view(u0, i).clone()
        t1, t2 = s.empty(), s.empty()
        s.ssprk3_steps(u, dts[i], steps, t1, t2)
        out.append(u)
    torch.cuda.synchronize()
    return out


def assert_bitwise(batch, arena, singles):
    bad = [i for i in range(batch.ncases) if not torch.equal(batch.view(arena, i), singles[i])]
    assert not bad, f"cases differing from their single runs: {bad}"


def test_small_ensemble_bit_identical(gpu):
    print("case batch small ensemble", flush=True)
    cases = ensemble.khi_ensemble(seed=11, ncases=12, grids=(8, 16), degrees=(1, 2, 3, 4, 5))
    batch, u0, dts = ensemble.khi_batch(cases)
    steps = 10
    singles = single_runs(batch, u0, dts, steps)
    for tuning in [dict(block=256), dict(block=128), dict(block=256, tile_elems=[3] * 8, grid=7)]:
        print(f"  tuning {tuning}", flush=True)
        batch.set_tuning(**tuning)
        u, t1, t2 = u0.clone(), batch.empty(), batch.empty()
        batch.ssprk3_steps(u, steps, t1, t2)
        torch.cuda.synchronize()
        assert_bitwise(batch, u, singles)


def test_batch_rhs_matches_oracle(gpu):
    print("case batch rhs vs oracle", flush=True)
    cases = ensemble.khi_ensemble(seed=5, ncases=4, grids=(8,), degrees=(2, 3, 4, 5))
    batch, u0, _ = ensemble.khi_batch(cases)
    du = batch.empty()
    batch.rhs(u0, du)
    torch.cuda.synchronize()
    for i, c in enumerate(cases):
        oc = O.Case(2, c.polydeg, c.cells, [-1, -1], [1, 1], volume=O.SHOCK_CAPTURING,
                    alpha_max=0.002, alpha_min=0.001)
        u_ref = batch.case(i).to_reference_layout(batch.view(u0, i))
        ref = oc.rhs(u_ref)
        got = batch.case(i).to_reference_layout(batch.view(du, i))
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-12


def test_full_ensemble_autotuned_bit_identical(gpu):
    """256 cases, grids from {8,16,32,64}^2, N from {2,3,4,5}, 100 steps."""
    print("case batch full ensemble", flush=True)
    cases = ensemble.khi_ensemble(seed=2024, ncases=256)
    batch, u0, dts = ensemble.khi_batch(cases)
    tuning = batch.autotune(u0, batch.empty(), reps=3)
    print(f"  autotuned {tuning}", flush=True)
    steps = 100
    u, t1, t2 = u0.clone(), batch.empty(), batch.empty()
    batch.ssprk3_steps(u, steps, t1, t2)
    torch.cuda.synchronize()
    assert torch.isfinite(u).all()
    singles = single_runs(batch, u0, dts, steps)
    assert_bitwise(batch, u, singles)


def test_batch_rejects_mixed_policies(gpu):
    cfgs = ensemble.khi_configs(ensemble.khi_ensemble(seed=1, ncases=2, grids=(8,)))
    cfgs[1].volume_integral = dg.SPLIT_FORM
    with pytest.raises(dg.DgsemError, match="one volume integral"):
        dg.Batch(cfgs)
