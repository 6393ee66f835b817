"""Layout conversion between the reference's solution views and the internal
layout.  The reference stores u(ie, je[, ke], dof, var) in a Kokkos::View
whose layout is the execution space's default (src/base/global_def.hpp:9-37):
row-major on host backends, column-major on GPU backends.  Both conversions
are pure permutations, so they must be bit-exact against numpy's own
row-major / column-major orderings of the same array."""
import numpy as np
import pytest
import torch

import dgsem_amd as dg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,N,cells", [(2, 3, (5, 3)), (3, 2, (4, 3, 2)), (3, 3, (2, 5, 3))])
def test_layout_left_and_right_round_trip(gpu, dim, N, cells):
    print(f"case layout dim={dim} N={N} cells={cells}", flush=True)
    cfg = dg.SolverConfig(dim=dim, polydeg=N, cells=cells, domain_left=[0.0] * dim,
                          domain_right=[1.0] * dim)
    s = dg.StructuredSolver(cfg)
    nn, nv = (N + 1) ** dim, dim + 2
    rng = np.random.default_rng(7)
    u5 = rng.standard_normal(tuple(cells) + (nn, nv))  # u(ie, je[, ke], dof, var)
    internal = s.from_reference_layout(u5.reshape(-1, nn, nv))
    want = np.transpose(u5.reshape(-1, nn, nv), (0, 2, 1)).ravel()
    assert np.array_equal(internal.cpu().numpy().ravel(), want)

    for layout, order in ((dg.LAYOUT_RIGHT, "C"), (dg.LAYOUT_LEFT, "F")):
        flat = torch.from_numpy(u5.ravel(order=order).copy()).cuda()
        got = s.pack_layout(flat, layout)
        assert torch.equal(got, internal), order
        back = s.unpack_layout(internal, layout)
        assert torch.equal(back, flat), order


def test_unknown_layout_is_rejected(gpu):
    cfg = dg.SolverConfig(dim=2, polydeg=3, cells=(2, 2), domain_left=[0.0, 0.0],
                          domain_right=[1.0, 1.0])
    s = dg.StructuredSolver(cfg)
    u = s.empty()
    with pytest.raises(dg.DgsemError, match="unknown layout"):
        s.pack_layout(u, 7)
