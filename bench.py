#!/usr/bin/env python3
"""Throughput benchmark of the DGSEM right-hand side on one MI355X.

Workload: 3D Euler Taylor-Green vortex on [0, 2pi]^3, 64^3 elements, N = 3
(16.8 M nodes x 5 conserved variables), split-form Chandrashekar flux
differencing, LLF surface flux, SSPRK3 -- the forward-throughput case the
metric is quoted on.  One "step" is one SSPRK3 step (three fused RHS + stage
kernels).  The state is resident in HBM before the timed region.

Prints ONE JSON line.  value = FP64 DOF-updates/s per SSPRK3 stage
= elements * (N+1)^3 / (step time / 3).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector / matrix peak
HBM_PEAK_GBS = 8000.0
# Algorithmic FLOP convention of the volume kernel (per pair / per node), 3D:
# a Chandrashekar flux with its averaged normal (85) plus the two-sided
# accumulation (22); 15 per node for the hoisted primitives.
C_PAIR_3D, C_NODE_3D = 107, 15


def tgv_state(coords: np.ndarray, gamma: float = 1.4) -> np.ndarray:
    """Taylor-Green vortex (Mach 0.1) in conservative variables, reference
    layout u[e][dof][var]."""
    from dgsem_amd import tgv
    return tgv.tgv_state(coords, gamma)


def volume_flops(E: int, N: int) -> float:
    n = N + 1
    pairs = 3 * n * n * N * (N + 1) // 2
    return float(E) * (pairs * C_PAIR_3D + n ** 3 * C_NODE_3D)


def stage_bytes(E: int, N: int, geometry: str, box_cells=None) -> float:
    """Algorithmic HBM bytes of one fused RHS + SSPRK3 stage launch, averaged
    over the three stage kinds: the state read once (stage 1), the state and
    the base state (stages 2 and 3), the result written once, and the metric
    terms (nodal: Ja and 1/J per node; affine: 10 doubles per element).

    box_cells: the cells per direction of a periodic box mesh whose nodal
    metric terms the line kernel reads from the box table.  Ja^r_c depends
    only on the element's two indices other than c, so the table holds, per
    direction c, one row of 3 x (N+1)^3 doubles for each pair of those
    indices; with 1/J still per node."""
    nodes = E * (N + 1) ** 3
    state = nodes * 5 * 8
    if geometry != "nodal":
        metrics = E * 10 * 8
    elif box_cells is not None:
        c = box_cells
        pairs = c[1] * c[2] + c[2] * c[0] + c[0] * c[1]
        metrics = pairs * 3 * (N + 1) ** 3 * 8 + nodes * 8
    else:
        metrics = nodes * (3 * 3 + 1) * 8
    return state * (1.0 + 2.0 / 3.0) + state + metrics


def read_profile(path: str) -> dict:
    """Figures of the fused kernel from a committed rocprofv3 summary: HBM
    bytes per dispatch (FETCH_SIZE + WRITE_SIZE, calibrated), the FP64 FLOPs
    per dispatch counted from the FP64 VALU instruction counters, and the
    kernel's average duration in the kernel trace."""
    if not os.path.exists(path):
        return {}
    text = open(path).read()

    def num(key):
        m = re.search(key + r":\s*([0-9.eE+-]+)", text)
        return float(m.group(1)) if m else None

    out = {}
    f, w = num("fetch_bytes_per_dispatch"), num("write_bytes_per_dispatch")
    if f is not None and w is not None:
        out["traffic"] = f + w
    if num("fp64_flop_per_dispatch") is not None:
        out["counter_flops"] = num("fp64_flop_per_dispatch")
    if num("kernel_avg_us") is not None:
        out["kernel_ms"] = num("kernel_avg_us") * 1e-3
    return out


def _host_description() -> dict:
    """CPU model, sockets and cores of this host (lscpu), and the cores this
    process may run on."""
    info = {"model": None, "sockets": None, "cores_per_socket": None, "threads_per_core": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, subprocess.SubprocessError):
        pass
    info["affinity_cpus"] = len(os.sched_getaffinity(0))
    return info


def _cpu_baseline_child(N: int, cells: int, steps: int, warmup: int) -> None:
    """Runs in a child process started by cpu_baseline() with the OpenMP
    environment set before the runtime starts; prints one JSON line."""
    from oracle import oracle as O

    threads = int(os.environ["OMP_NUM_THREADS"])
    c = O.Case(3, N, (cells,) * 3, [0, 0, 0], [2 * math.pi] * 3, threads=threads)
    u = c.init_tgv()
    dt = 1e-4
    for _ in range(warmup):
        u = c.ssprk3(u, dt, 1)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        u = c.ssprk3(u, dt, 1)
        times.append(time.perf_counter() - t0)
    print(json.dumps({"times": times, "dof": c.E * c.ndofs}))


def cpu_baseline(N: int, cells: int, steps: int, warmup: int) -> dict:
    """The CPU restatement of the reference (OpenMP over elements, one thread
    per core of this process's CPU share, OMP_PLACES=cores,
    OMP_PROC_BIND=close) on the same workload: `warmup` untimed SSPRK3 steps,
    then the median of `steps` steps.  It runs in a child process so that the
    OpenMP placement variables take effect before the runtime starts."""
    host = _host_description()
    threads = int(os.environ.get("OMP_NUM_THREADS", host["affinity_cpus"]))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores", OMP_PROC_BIND="close")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-child",
           "--polydeg", str(N), "--cpu-cells", str(cells), "--cpu-steps", str(steps),
           "--cpu-warmup", str(warmup)]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed: {r.stderr[-2000:]}")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    t_step = statistics.median(res["times"])
    dof = res["dof"]
    return {"value": dof / (t_step / 3.0), "unit": "DOF-updates/s per SSPRK3 stage",
            "cores": threads, "kind": "port",
            "sample": f"TGV N={N} on {cells}^3 elements ({dof} nodes), median of {steps} "
                      f"SSPRK3 steps after {warmup} warm-up steps, oracle/ restatement, "
                      f"OpenMP {threads} threads, OMP_PLACES=cores OMP_PROC_BIND=close",
            "ms_per_step": t_step * 1e3, "step_ms": [t * 1e3 for t in res["times"]],
            "host": host}


def batch_measure(ncases: int = 256, steps: int = 100, seed: int = 2024, sequential: bool = True) -> dict:
    """Ensemble of 2D Kelvin-Helmholtz cases (grids from {8,16,32,64}^2, N from
    {2,3,4,5}): `steps` SSPRK3 steps of the autotuned batch (graph replay of
    one persistent launch per stage) against the same steps run case after
    case with the single-case handles; per-case results compared bitwise.
    sequential=False times the batch alone (for profiling it).

    Algorithmic HBM bytes of one stage launch, per node: the state (4
    doubles) read once at stage 1 and with the base state at stages 2 and 3
    (1 + 2/3 reads on average), the result written once, and the nodal
    metric terms Ja (2 x 2) and 1/J."""
    import torch

    from dgsem_amd import ensemble

    cases = ensemble.khi_ensemble(seed=seed, ncases=ncases)
    batch, u0, dts = ensemble.khi_batch(cases)
    nodes = sum(batch.case(i).n_elements * batch.case(i).ndofs for i in range(ncases))
    tuning = batch.autotune(u0, batch.empty(), reps=3)
    ub, t1, t2 = u0.clone(), batch.empty(), batch.empty()
    batch.ssprk3_steps(ub, 1, t1, t2)  # graph capture and warm-up
    ub.copy_(u0)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    batch.ssprk3_steps(ub, steps, t1, t2)
    ev1.record()
    torch.cuda.synchronize()
    t_batch = ev0.elapsed_time(ev1)
    stage_bytes_ = nodes * (4 * 8 * (1.0 + 2.0 / 3.0 + 1.0) + 5 * 8)
    stage_ms = t_batch / (3 * steps)
    res = {"workload": f"khi2d_ensemble_{ncases}_cases_{steps}_steps_seed{seed}",
           "nodes": nodes, "batched_ms": t_batch, "stage_ms": stage_ms,
           "bytes_per_stage": stage_bytes_,
           "hbm": {"achieved": stage_bytes_ / (stage_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": stage_bytes_ / (stage_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "dof_updates_per_s": nodes * 3 * steps / (t_batch * 1e-3), "tuning": tuning}
    if not sequential:
        return res
    singles = []
    for i in range(ncases):
        s = batch.case(i)
        singles.append((s, batch.view(u0, i).clone(), s.empty(), s.empty()))
    s, u, a, b = singles[0]
    s.ssprk3_steps(u.clone(), dts[0], 1, a, b)  # warm-up
    torch.cuda.synchronize()
    ev0.record()
    for i, (s, u, a, b) in enumerate(singles):
        s.ssprk3_steps(u, dts[i], steps, a, b)
    ev1.record()
    torch.cuda.synchronize()
    t_seq = ev0.elapsed_time(ev1)
    identical = all(torch.equal(batch.view(ub, i), singles[i][1]) for i in range(ncases))
    res.update({"sequential_ms": t_seq, "speedup": t_seq / t_batch, "bit_identical": identical})
    return res


def adjoint_measure(cells: int = 16, N: int = 3, steps: int = 200, every: int = 20) -> dict:
    """3D Navier-Stokes TGV (Re 1600, Pr 0.72): time of the forward trajectory
    with the kinetic-energy objective against the checkpointed reverse sweep
    (recompute between checkpoints + adjoint steps)."""
    import torch

    import dgsem_amd as dg
    from dgsem_amd import adjoint, tgv

    cfg = dg.SolverConfig(dim=3, polydeg=N, cells=(cells,) * 3, domain_left=(0, 0, 0),
                          domain_right=(2 * math.pi,) * 3, equation=dg.NAVIER_STOKES,
                          mu=6.25e-4, prandtl=0.72)
    s = dg.StructuredSolver(cfg)
    u0 = s.from_reference_layout(tgv.tgv_state(s.node_coordinates()))
    dt = tgv.tgv_dt(cells, N)
    adjoint.kinetic_energy_objective(s, u0, dt, 4, 2)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J, ckpt = adjoint.kinetic_energy_trajectory(s, u0, dt, steps, every)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    adjoint.kinetic_energy_gradient(s, ckpt, dt, steps, every)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return {"workload": f"tgv3d_navier_stokes_{cells}^3_N{N}_{steps}_steps_checkpoint_{every}",
            "forward_ms_per_step": (t1 - t0) / steps * 1e3,
            "adjoint_ms_per_step": (t2 - t1) / steps * 1e3,
            "ratio": (t2 - t1) / (t1 - t0), "objective": J}


def vtu_measure(cells: int = 32, N: int = 3) -> dict:
    """VTU output of a 3D TGV state (dgsem_write_vtu): wall time of one file
    (GPU reorder + copy + write to a local temporary file) and its size."""
    import tempfile

    import torch

    import dgsem_amd as dg

    cfg = dg.SolverConfig(dim=3, polydeg=N, cells=(cells,) * 3, domain_left=(0, 0, 0),
                          domain_right=(2 * math.pi,) * 3)
    s = dg.StructuredSolver(cfg)
    u = s.from_reference_layout(tgv_state(s.node_coordinates()))
    torch.cuda.synchronize()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tgv.vtu")
        s.write_vtu(u, path)  # warm-up (allocations, page cache)
        t0 = time.perf_counter()
        s.write_vtu(u, path)
        t = time.perf_counter() - t0
        size = os.path.getsize(path)
    return {"workload": f"tgv3d_{cells}^3_N{N}_vtu", "points": s.n_elements * s.ndofs,
            "file_bytes": size, "ms": t * 1e3, "GB_per_s": size / t / 1e9}


def geometry_measure(cells: int = 32, N: int = 3) -> dict:
    """Element geometry of a warped 3D mesh from mapped node coordinates
    (dgsem_set_node_coordinates: GPU kernel plus the host mirrors' copies)
    against the host construction of the linear mesh at creation."""
    import torch

    import dgsem_amd as dg

    cfg = dg.SolverConfig(dim=3, polydeg=N, cells=(cells,) * 3, domain_left=(-1, -1, -1),
                          domain_right=(1, 1, 1), volume_integral=dg.SHOCK_CAPTURING)
    t0 = time.perf_counter()
    s = dg.StructuredSolver(cfg)
    t_create = time.perf_counter() - t0
    ref = s.reference_coordinates()
    warp = np.prod(np.sin(np.pi * ref), axis=-1)[..., None]
    coords = torch.from_numpy(ref + 0.1 * warp).cuda()
    s.set_node_coordinates(coords)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.set_node_coordinates(coords)
    t_dev = time.perf_counter() - t0
    return {"workload": f"warped3d_{cells}^3_N{N}", "elements": s.n_elements,
            "host_create_ms": t_create * 1e3, "device_geometry_ms": t_dev * 1e3}


def _step_rate(solver, u, dt: float, steps: int, warmup: int):
    """(ms per SSPRK3 step, DOF-updates/s per stage, ms per stage launch) of
    `steps` timed steps after `warmup`, the state resident on the GPU."""
    import torch

    import dgsem_amd as dg

    rk = dg.SSPRK3(solver)
    for _ in range(warmup):
        rk.step(u, dt)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(steps):
        rk.step(u, dt)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    return ms, solver.n_elements * solver.ndofs / (ms / 3.0 * 1e-3), ms / 3.0


def tgv_high_order_measure(cells: int = 32, N: int = 7, steps: int = 10) -> dict:
    """The Taylor-Green vortex at N = 7 on 32^3 elements (the same 16.8 M
    nodes as the headline workload): step rate and the fused kernel's volume
    FLOP rate against FP64 peak."""
    import dgsem_amd as dg

    cfg = dg.SolverConfig(dim=3, polydeg=N, cells=(cells,) * 3, domain_left=(0, 0, 0),
                          domain_right=(2 * math.pi,) * 3)
    s = dg.StructuredSolver(cfg)
    u = s.from_reference_layout(tgv_state(s.node_coordinates()))
    dt = 0.5 * (2 * math.pi / cells) / (3 * (2 * N + 1) * (1.0 + math.sqrt(1.4 * (100.0 / 1.4 + 0.25))))
    ms, rate, stage_ms = _step_rate(s, u, dt, steps, 2)
    tflops = volume_flops(s.n_elements, N) / (stage_ms * 1e-3) / 1e12
    return {"workload": f"tgv3d_euler_{cells}^3_N{N}_splitform_llf_ssprk3",
            "kernel": "rhs_node_kernel" if N >= 7 else "rhs_kernel",
            "nodes": s.n_elements * s.ndofs, "ms_per_step": ms, "dof_updates_per_s": rate,
            "volume_tflops": tflops, "frac_fp64_peak": tflops / FP64_PEAK_TFLOPS}


def ns_high_order_measure(cells: int = 32, N: int = 7, steps: int = 10, mu: float = 6.25e-4) -> dict:
    """The Navier-Stokes Taylor-Green vortex (Re = 1600 at mu = 6.25e-4,
    Pr = 0.72) at N = 7 on 32^3 elements: step rate of the BR1 path whose
    gradient, viscous fluxes and sums follow the reference's order and
    roundings (two node-parallel BR1 kernels and the node RHS kernel per
    stage)."""
    import dgsem_amd as dg

    cfg = dg.SolverConfig(dim=3, polydeg=N, cells=(cells,) * 3, domain_left=(0, 0, 0),
                          domain_right=(2 * math.pi,) * 3, equation=dg.NAVIER_STOKES, mu=mu,
                          prandtl=0.72)
    s = dg.StructuredSolver(cfg)
    u = s.from_reference_layout(tgv_state(s.node_coordinates()))
    dt = 0.5 * (2 * math.pi / cells) / (3 * (2 * N + 1) * (1.0 + math.sqrt(1.4 * (100.0 / 1.4 + 0.25))))
    ms, rate, stage_ms = _step_rate(s, u, dt, steps, 2)
    return {"workload": f"tgv3d_navier_stokes_{cells}^3_N{N}_mu{mu}_splitform_llf_br1_ssprk3",
            "nodes": s.n_elements * s.ndofs, "ms_per_step": ms, "dof_updates_per_s": rate}


def sedov_measure(cells: int = 32, N: int = 3, steps: int = 20) -> dict:
    """The 3D Sedov blast with subcell shock capturing (alpha_max 0.5, HLLC
    subcell flux, positivity floor 1e-13) on 32^3 elements at N = 3: step
    rate and the share of elements the indicator blends."""
    import dgsem_amd as dg
    from dgsem_amd import sedov

    cfg = dg.SolverConfig(dim=3, polydeg=N, cells=(cells,) * 3, domain_left=(-1, -1, -1),
                          domain_right=(1, 1, 1), volume_integral=dg.SHOCK_CAPTURING,
                          fv_flux=dg.FLUX_HLLC, alpha_max=0.5, alpha_min=0.001,
                          positivity_floor=1e-13)
    s = dg.StructuredSolver(cfg)
    u0 = sedov.sedov_state(s.node_coordinates())
    dt = sedov.sedov_dt(u0, N, 2.0 / cells)
    u = s.from_reference_layout(u0)
    ms, rate, _ = _step_rate(s, u, dt, steps, 2)
    troubled = float((s.alpha() > 0).double().mean())
    return {"workload": f"sedov3d_{cells}^3_N{N}_shock_capturing_hllc_fv_ssprk3",
            "nodes": s.n_elements * s.ndofs, "ms_per_step": ms, "dof_updates_per_s": rate,
            "troubled_fraction": troubled}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--polydeg", type=int, default=3)
    ap.add_argument("--geometry", choices=["nodal", "affine"], default="nodal")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cells", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--cpu-warmup", type=int, default=2)
    ap.add_argument("--cpu-baseline-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--settle-s", type=float, default=0.5,
                    help="after the warm-up steps, keep stepping untimed for this many "
                         "seconds so that the timed window runs at steady clocks")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-ensemble and adjoint measurements")
    ap.add_argument("--batch-only", type=int, default=0, metavar="STEPS",
                    help="time only the batched ensemble for STEPS SSPRK3 steps (no sequential "
                         "comparison) and print its figures")
    ap.add_argument("--sedov-only", type=int, default=0, metavar="STEPS",
                    help="time only the 3D Sedov blast (32^3, N = 3, shock capturing with HLLC "
                         "subcells) for STEPS SSPRK3 steps and print its figures")
    ap.add_argument("--ns-n7", action="store_true",
                    help="also time the Navier-Stokes TGV at N = 7 on 32^3 elements")
    ap.add_argument("--profile-copy-mib", type=int, default=0,
                    help="after the run, launch an elementwise copy kernel of this size "
                         "(calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()

    if args.cpu_baseline_child:
        _cpu_baseline_child(args.polydeg, args.cpu_cells, args.cpu_steps, args.cpu_warmup)
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(the RHS would shard by element blocks with a face-trace halo exchange). "
              "Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    import dgsem_amd as dg

    if args.batch_only > 0 or args.sedov_only > 0:
        if args.batch_only > 0:
            res = {"batch": batch_measure(steps=args.batch_only, sequential=False)}
        else:
            res = {"sedov": sedov_measure(steps=args.sedov_only)}
        if args.profile_copy_mib > 0:
            a = torch.ones(args.profile_copy_mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
            b = torch.empty_like(a)
            for _ in range(3):
                torch.mul(a, 1.0, out=b)
            torch.cuda.synchronize()
            res["copy_calibration_bytes"] = a.numel() * 8
        print(json.dumps(res))
        return 0

    N, cells = args.polydeg, args.cells
    geometry = dg.GEOMETRY_NODAL if args.geometry == "nodal" else dg.GEOMETRY_AFFINE
    cfg = dg.SolverConfig(dim=3, polydeg=N, cells=(cells,) * 3, domain_left=(0, 0, 0),
                          domain_right=(2 * math.pi,) * 3, geometry=geometry)
    solver = dg.StructuredSolver(cfg)
    u = solver.from_reference_layout(tgv_state(solver.node_coordinates()))
    rk = dg.SSPRK3(solver)
    E, nn = solver.n_elements, solver.ndofs
    # dt = CFL dx / (3 (2N+1) lambda_max), CFL 0.5, lambda_max = max |v| + c
    lam = 1.0 + math.sqrt(1.4 * (100.0 / 1.4 + 0.25) / 1.0)
    dt = 0.5 * (2 * math.pi / cells) / (3 * (2 * N + 1) * lam)

    # Untimed warm-up: the W steps asked for, then further steps until the
    # GPU has run stage launches for at least --settle-s seconds, so that the
    # timed window sees the clocks the kernel sustains rather than the boost
    # of a cold start.
    for _ in range(args.warmup):
        rk.step(u, dt)
    torch.cuda.synchronize()
    settle_steps, t_settle = 0, time.perf_counter()
    while time.perf_counter() - t_settle < args.settle_s:
        for _ in range(10):
            rk.step(u, dt)
        settle_steps += 10
        torch.cuda.synchronize()

    # Timed window: K steps, timed by the host clock and by HIP events on the
    # stream the kernels run on (torch's current stream, which the solver
    # handle was created with).  Every launch in the window is the fused
    # RHS + stage kernel, so the event time over 3 K launches is that
    # kernel's average launch duration in the same window as ms_per_step.
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        rk.step(u, dt)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if not bool(torch.isfinite(u).all()):
        print("non-finite state after the timed steps", file=sys.stderr)
        return 1
    ms_per_step = wall / args.steps * 1e3
    value = E * nn / (ms_per_step / 3.0 * 1e-3)
    t_kernel_ms = ev0.elapsed_time(ev1) / (3 * args.steps)

    flops = volume_flops(E, N)
    # the line kernels (N < 7) read a box mesh's nodal metric terms from the
    # box table; the node kernel (N = 7) reads the per-node arrays
    hbm_bytes = stage_bytes(E, N, args.geometry,
                            (cells,) * 3 if args.geometry == "nodal" and N < 7 else None)
    intensity = flops / hbm_bytes
    ridge = FP64_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9)
    bound = "hbm" if intensity < ridge else "fp64"
    prof = read_profile(os.path.join(ROOT, "profiles", f"rhs_stage_tgv{cells}_n{N}_{args.geometry}_pmc.txt"))
    traffic = prof.get("traffic")
    achieved_gbs = hbm_bytes / (t_kernel_ms * 1e-3) / 1e9
    achieved_tf = flops / (t_kernel_ms * 1e-3) / 1e12
    kernel_name = ("rhs_node_kernel (node-parallel fused volume+surface+stage)" if N >= 7
                   else "rhs_kernel (fused volume+surface+stage)")
    if bound == "hbm":
        roof = {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved_gbs / HBM_PEAK_GBS}
    else:
        roof = {"bound": "fp64", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved_tf / FP64_PEAK_TFLOPS}
    roof.update({
        "traffic": traffic, "kernel": kernel_name, "kernel_ms": t_kernel_ms,
        "bytes_per_launch": hbm_bytes, "flop_per_launch": flops,
        "arithmetic_intensity": intensity, "ridge_flop_per_byte": ridge,
        # the other roof, side by side: algorithmic volume FLOPs against FP64 peak
        "fp64": {"achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": achieved_tf / FP64_PEAK_TFLOPS,
                 # FP64 VALU instructions counted by rocprofv3 (per dispatch, x 64 lanes,
                 # FMA as 2) over this window's launch duration
                 "counter_flop_per_launch": prof.get("counter_flops"),
                 "counter_achieved": (None if prof.get("counter_flops") is None
                                      else prof["counter_flops"] / (t_kernel_ms * 1e-3) / 1e12)},
        "measured_GB_per_s": None if traffic is None else traffic / (t_kernel_ms * 1e-3) / 1e9,
        # average duration of the same kernel in the committed rocprofv3 trace
        "profile_kernel_ms": prof.get("kernel_ms"),
    })

    out = {
        "metric": "FP64 DOF-updates/s per SSPRK3 stage",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Taylor-Green vortex initial state)",
        "config": {"workload": f"tgv3d_euler_{cells}^3_N{N}_splitform_llf_ssprk3",
                   "elements": E, "polydeg": N, "nodes": E * nn, "nvars": 5,
                   "geometry": args.geometry, "dt": dt},
        "settle": {"steps": settle_steps, "min_s": args.settle_s},
        "dof_var_updates_per_s": value * 5,
        "adjoint_forward_step_ratio": None,
        "roofline": roof,
    }
    if not args.no_extras:
        out["batch"] = batch_measure()
        out["adjoint"] = adjoint_measure()
        out["adjoint_forward_step_ratio"] = out["adjoint"]["ratio"]
        out["vtu"] = vtu_measure()
        out["geometry"] = geometry_measure()
        out["tgv_n7"] = tgv_high_order_measure()
        out["sedov"] = sedov_measure()
    if args.ns_n7:
        out["ns_n7"] = ns_high_order_measure()
    if args.profile_copy_mib > 0:
        a = torch.ones(args.profile_copy_mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = a.numel() * 8
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(N, args.cpu_cells, args.cpu_steps, args.cpu_warmup)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
