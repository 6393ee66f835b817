"""Simulation observers for SSPRK3.solve, mirroring the reference's
observer set (src/observer): PrintObserver (print_observer.cpp:10-20),
VTUOutputObserver (vtu_output_observer.hpp:18-233), and AnalysisObserver with
the StopOnNaN and VolumeAverageCsvWriter handlers (analysis_observer.hpp:
56-186).  The analysis runs on the GPU (dgsem_analyze) and the VTU file is
written by dgsem_write_vtu; these classes only decide when.

Protocol (integrator_base.hpp:21-40): before every step the integrator calls
on_step(step, time); an observer returns True to stop the run (the
reference's has_nan flag).  After the loop it calls on_finish(step, time).
"""
from __future__ import annotations

import sys


class SimulationObserver:
    def on_step(self, step: int, time: float) -> bool:
        return False

    def on_finish(self, step: int, time: float) -> None:
        pass


class PrintObserver(SimulationObserver):
    def __init__(self, interval: int, out=None):
        self.interval = interval
        self.out = out or sys.stdout

    def on_step(self, step, time):
        if step % self.interval == 0:
            print(f"Iter: {step:5d}  Time: {time:.8e}", file=self.out)
        return False

    def on_finish(self, step, time):
        print(f"Finished at iter {step}, time {time}", file=self.out)


class VTUOutputObserver(SimulationObserver):
    """Writes <output_path>_step<N>.vtu at step 0, every output_interval
    steps (if positive) and at the end unless that step was just written."""

    def __init__(self, output_path: str, solver, u, output_interval: int = -1):
        self.output_path = output_path
        self.solver = solver
        self.u = u
        self.output_interval = output_interval
        self.last_written_step = -1
        self.written: list[str] = []

    def filename(self, step: int) -> str:
        return f"{self.output_path}_step{step}.vtu"

    def _write(self, step):
        path = self.filename(step)
        self.solver.write_vtu(self.u, path)
        self.written.append(path)
        self.last_written_step = step

    def on_step(self, step, time):
        if step == 0:
            self._write(step)
        if self.output_interval >= 0 and step % self.output_interval == 0 and step != self.last_written_step:
            self._write(step)
        return False

    def on_finish(self, step, time):
        if step != self.last_written_step:
            self._write(step)


def stop_on_nan(step, time, stats, out=None) -> bool:
    """StopOnNaN: stop when any node holds a non-finite value."""
    if stats["nonfinite_nodes"] > 0:
        print(f"NaN detected at iter {step}, time {time}", file=out or sys.stderr)
        return True
    return False


class VolumeAverageCsvWriter:
    """step,time,volume_average_kinetic_energy,volume_average_entropy rows."""

    def __init__(self, output_path: str):
        self.output_path = output_path
        self.f = open(output_path, "w")
        self.f.write("step,time,volume_average_kinetic_energy,volume_average_entropy\n")

    def __call__(self, step, time, stats) -> bool:
        self.f.write(f"{step},{time:.16e},{stats['mean_kinetic_energy']:.16e},"
                     f"{stats['mean_entropy']:.16e}\n")
        self.f.flush()
        return False

    def close(self):
        self.f.close()


class AnalysisObserver(SimulationObserver):
    """Runs the on-device analysis of u and passes its results to each
    handler(step, time, stats); any handler returning True stops the run."""

    def __init__(self, solver, u, *handlers, analyze_on_finish: bool = False):
        self.solver = solver
        self.u = u
        self.handlers = handlers
        self.analyze_on_finish = analyze_on_finish

    def _run(self, step, time) -> bool:
        stats = self.solver.analyze(self.u)
        stop = False
        for h in self.handlers:
            stop = bool(h(step, time, stats)) or stop
        return stop

    def on_step(self, step, time):
        return self._run(step, time)

    def on_finish(self, step, time):
        if self.analyze_on_finish:
            self._run(step, time)
