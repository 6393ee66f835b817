"""The observer protocol of SSPRK3.solve (integrator_base.hpp:21-40) and the
observers' schedules, with a stand-in solver (no GPU): observers run before
every step with (iteration, time), a True return stops the run after that
step, on_finish runs once at the end; the VTU observer writes at step 0,
every interval and at the end; the CSV writer's rows."""
import numpy as np

from dgsem_amd.observers import (AnalysisObserver, PrintObserver, SimulationObserver,
                                 VolumeAverageCsvWriter, VTUOutputObserver, stop_on_nan)
from dgsem_amd.solver import SSPRK3


class FakeSolver:
    def __init__(self, nan_after=None):
        self.stages = []
        self.files = []
        self.nan_after = nan_after
        self.calls = 0

    def empty(self):
        return np.zeros(1)

    def ssprk3_stage(self, stage, dt, u_base, u_k, u_out):
        self.stages.append(stage)

    def write_vtu(self, u, path):
        self.files.append(path)

    def analyze(self, u):
        self.calls += 1
        bad = 1 if self.nan_after is not None and self.calls > self.nan_after else 0
        return {"nonfinite_nodes": bad, "mean_kinetic_energy": 0.5 * self.calls,
                "mean_entropy": -float(self.calls)}


class Recorder(SimulationObserver):
    def __init__(self):
        self.steps, self.finish = [], None

    def on_step(self, step, time):
        self.steps.append((step, time))
        return False

    def on_finish(self, step, time):
        self.finish = (step, time)


def test_solve_calls_observers_before_each_step():
    s = FakeSolver()
    rk = SSPRK3(s, final_time=0.45)
    rec = Recorder()
    rk.add_observer(rec)
    rk.solve(np.zeros(1), 0.1)
    assert [k for k, _ in rec.steps] == [0, 1, 2, 3, 4]
    assert np.allclose([t for _, t in rec.steps], [0.0, 0.1, 0.2, 0.3, 0.4])
    assert rec.finish[0] == 5 and s.stages == [1, 2, 3] * 5


def test_vtu_schedule():
    s = FakeSolver()
    rk = SSPRK3(s, final_time=0.45)
    rk.add_observer(VTUOutputObserver("out/run", s, np.zeros(1), output_interval=2))
    rk.solve(np.zeros(1), 0.1)
    assert s.files == ["out/run_step0.vtu", "out/run_step2.vtu", "out/run_step4.vtu",
                       "out/run_step5.vtu"]
    s2 = FakeSolver()
    rk2 = SSPRK3(s2, final_time=0.25)
    rk2.add_observer(VTUOutputObserver("x", s2, np.zeros(1)))
    rk2.solve(np.zeros(1), 0.1)
    assert s2.files == ["x_step0.vtu", "x_step3.vtu"]


def test_stop_on_nan_and_csv(tmp_path):
    s = FakeSolver(nan_after=2)
    rk = SSPRK3(s, final_time=1.0)
    csv = VolumeAverageCsvWriter(str(tmp_path / "stats.csv"))
    rk.add_observer(AnalysisObserver(s, np.zeros(1), lambda *a: stop_on_nan(*a, out=open("/dev/null", "w")), csv))
    rk.add_observer(PrintObserver(100, out=open("/dev/null", "w")))
    rk.solve(np.zeros(1), 0.1)
    csv.close()
    # analyses before steps 0, 1, 2; the third reports a NaN, so step 2 runs and the loop ends
    assert rk.iter == 3 and s.stages == [1, 2, 3] * 3
    rows = open(tmp_path / "stats.csv").read().splitlines()
    assert rows[0] == "step,time,volume_average_kinetic_energy,volume_average_entropy"
    assert [r.split(",")[0] for r in rows[1:]] == ["0", "1", "2"]
    assert float(rows[2].split(",")[2]) == 1.0
