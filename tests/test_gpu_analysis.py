"""On-device analysis (dgsem_analyze) against the oracle's orc_analyze on
the same states: volume, kinetic-energy and entropy integrals (relative
1e-13: both sum the same terms, in different orders), non-finite count and
largest signal speed (a max of per-node values that agree to a few ulp:
the GPU fuses multiply-adds, the oracle does not)."""
import math

import numpy as np
import pytest
import torch

import dgsem_amd as dg
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,N,cells,geometry", [(3, 3, (4, 3, 5), dg.GEOMETRY_NODAL),
                                                  (3, 5, (2, 2, 2), dg.GEOMETRY_AFFINE),
                                                  (2, 4, (8, 6), dg.GEOMETRY_NODAL)])
def test_analysis_matches_oracle(gpu, dim, N, cells, geometry):
    print(f"case analysis dim={dim} N={N} cells={cells}", flush=True)
    left, right = ([0.0] * dim, [2 * math.pi] * dim) if dim == 3 else ([-1.0, -1.0], [1.0, 1.0])
    oc = O.Case(dim, N, cells, left, right)
    u = oc.init_tgv() if dim == 3 else oc.init_khi()
    s = dg.StructuredSolver(dg.SolverConfig(dim=dim, polydeg=N, cells=cells, domain_left=left,
                                            domain_right=right, geometry=geometry))
    got = s.analyze(s.from_reference_layout(u))
    ref = oc.analyze(u)
    if geometry == dg.GEOMETRY_AFFINE:
        # exact per-element metrics: the volume weights differ from the
        # reference's per-node values by their rounding noise
        tol = 1e-11
    else:
        tol = 1e-13
    for k, key in enumerate(["volume", "kinetic_energy", "entropy"]):
        assert abs(got[key] - ref[k]) <= tol * abs(ref[k]), key
    assert got["nonfinite_nodes"] == ref[3] == 0
    assert abs(got["max_speed"] - ref[4]) <= 1e-14 * ref[4]


def test_analysis_counts_non_finite_nodes(gpu):
    oc = O.Case(2, 3, (4, 4), [-1, -1], [1, 1])
    u = oc.init_khi()
    u[3, 5, 1] = np.nan
    u[7, 0, 3] = np.inf
    s = dg.StructuredSolver(dg.SolverConfig(dim=2, polydeg=3, cells=(4, 4), domain_left=[-1, -1],
                                            domain_right=[1, 1]))
    got = s.analyze(s.from_reference_layout(u))
    ref = oc.analyze(u)
    assert got["nonfinite_nodes"] == ref[3] == 2
    assert abs(got["kinetic_energy"] - ref[1]) <= 1e-13 * abs(ref[1])
