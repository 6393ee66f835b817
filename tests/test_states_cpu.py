"""The product-side initial states (dgsem_amd.tgv, dgsem_amd.sedov) equal the
oracle's at the same node coordinates."""
import numpy as np

from dgsem_amd import sedov, tgv
from oracle import oracle as O


def test_sedov_state_matches_oracle():
    c = O.Case(3, 3, (4, 4, 4), [-1.0] * 3, [1.0] * 3)
    got = sedov.sedov_state(c.coordinates())
    ref = c.init_sedov()
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0)
    # the blast sits at the origin over the 1e-5 background
    p = 0.4 * got[..., 4]
    assert p.min() >= 0.999999e-5 and p.max() > 100
    dt = sedov.sedov_dt(got, 3, 0.5)
    assert 0 < dt < 1e-3


def test_tgv_state_matches_oracle():
    c = O.Case(3, 3, (3, 3, 3), [0.0] * 3, [2 * np.pi] * 3)
    np.testing.assert_allclose(tgv.tgv_state(c.coordinates()), c.init_tgv(), rtol=1e-14, atol=1e-14)
