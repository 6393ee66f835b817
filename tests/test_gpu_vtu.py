"""GPU VTU output (dgsem_write_vtu) against the VTU oracle: the file the
library writes holds exactly the arrays the reference's VTUOutputObserver
would (cell order, VTK Lagrange point order, point data, cells), read back
with the oracle's reader.  Conservative variables, points and cells are
copies and compared bit for bit; the pressure is recomputed and compared
to 1e-14 relative (contracted multiply-adds on the GPU)."""
import numpy as np
import pytest

import dgsem_amd as dg
from oracle import oracle as O
from oracle import vtu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,N,cells", [(2, 3, (5, 4)), (2, 1, (3, 2)), (2, 7, (2, 3)),
                                         (3, 2, (3, 2, 4)), (3, 3, (2, 3, 2)), (3, 5, (2, 1, 2))])
def test_vtu_file_matches_oracle(gpu, tmp_path, dim, N, cells):
    print(f"case vtu dim={dim} N={N} cells={cells}", flush=True)
    left, right = [-1.0, -0.5, 0.0][:dim], [1.0, 0.75, 2.0][:dim]
    cfg = dg.SolverConfig(dim=dim, polydeg=N, cells=cells, domain_left=left, domain_right=right)
    gs = dg.StructuredSolver(cfg)
    oc = O.Case(dim, N, cells, left, right)
    rng = np.random.default_rng(N)
    u = rng.random(oc.shape) + 1.0
    u[..., dim + 1] += 20.0  # positive pressure
    path = tmp_path / "state.vtu"
    gs.write_vtu(gs.from_reference_layout(u), str(path))
    got = vtu.read_vtu(path)
    want = vtu.expected_arrays(dim, N, cells, gs.node_coordinates(), u)
    assert got["version"] == "2.2"
    assert got["points_count"] == u.shape[0] * u.shape[1] and got["cells_count"] == u.shape[0]
    names = list(want["point_data"])
    assert list(got["point_data"]) == names
    for k in names[:-1]:
        assert np.array_equal(got["point_data"][k], want["point_data"][k]), k
    p_got, p_want = got["point_data"]["pressure"], want["point_data"]["pressure"]
    assert np.min(p_want) > 0.0
    assert np.max(np.abs(p_got - p_want) / np.abs(p_want)) <= 1e-14
    assert np.array_equal(got["points"], want["points"])
    for k in ("connectivity", "offsets", "types"):
        assert np.array_equal(got[k], want[k]), k


def test_observers_drive_vtu_and_analysis(gpu, tmp_path):
    """SSPRK3.solve with a VTU observer every 2 steps and an analysis
    observer (StopOnNaN, CSV of volume averages): the files appear on the
    schedule and the last one holds the final state."""
    from dgsem_amd.observers import AnalysisObserver, VolumeAverageCsvWriter, VTUOutputObserver, stop_on_nan
    print("case vtu observers", flush=True)
    N, cells = 3, (8, 8)
    cfg = dg.SolverConfig(dim=2, polydeg=N, cells=cells, domain_left=[-1, -1], domain_right=[1, 1],
                          volume_integral=dg.SHOCK_CAPTURING, alpha_max=0.002)
    gs = dg.StructuredSolver(cfg)
    oc = O.Case(2, N, cells, [-1, -1], [1, 1])
    u = gs.from_reference_layout(oc.init_khi())
    rk = dg.SSPRK3(gs, final_time=5 * 1e-3 - 5e-4)
    vtu_obs = VTUOutputObserver(str(tmp_path / "khi"), gs, u, output_interval=2)
    csv = VolumeAverageCsvWriter(str(tmp_path / "stats.csv"))
    rk.add_observer(vtu_obs)
    rk.add_observer(AnalysisObserver(gs, u, stop_on_nan, csv))
    rk.solve(u, 1e-3)
    csv.close()
    assert [p.rsplit("_", 1)[1] for p in vtu_obs.written] == ["step0.vtu", "step2.vtu", "step4.vtu",
                                                             "step5.vtu"]
    last = vtu.read_vtu(vtu_obs.written[-1])
    want = vtu.expected_arrays(2, N, cells, gs.node_coordinates(), gs.to_reference_layout(u))
    assert np.array_equal(last["point_data"]["rho"], want["point_data"]["rho"])
    rows = open(tmp_path / "stats.csv").read().splitlines()
    assert len(rows) == 1 + 5
