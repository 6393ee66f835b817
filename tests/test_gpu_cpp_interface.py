"""The reference-style C++ interface (dgsem_amd/include/dgsem_hip.hpp): a
program composed like the reference's KHI example runs SSPRK3 on the GPU and
matches the restatement of the reference after the same steps."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "dgsem_amd", "examples", "khi_shock_capturing")


@pytest.mark.gpu
def test_cpp_khi_example_matches_oracle(gpu, tmp_path):
    assert os.path.exists(EXE), "build the library first (dgsem_amd/build.py)"
    nx, ny, steps, dt, amax = 16, 12, 30, 2e-4, 0.002
    out = tmp_path / "u.bin"
    print(f"case cpp khi {nx}x{ny} steps={steps}", flush=True)
    subprocess.run([EXE, str(nx), str(ny), str(steps), str(dt), str(amax), str(out)],
                   check=True, timeout=60)
    c = O.Case(2, 3, (nx, ny), [-1, -1], [1, 1], volume=O.SHOCK_CAPTURING, alpha_max=amax,
               alpha_min=0.001)
    ref = c.ssprk3(c.init_khi(), dt, steps)
    got = np.fromfile(out, dtype=np.float64).reshape(ref.shape)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-10


def run_example(name, tmp_path, *args):
    exe = os.path.join(ROOT, "dgsem_amd", "examples", name)
    assert os.path.exists(exe), "build the library first (dgsem_amd/build.py)"
    prefix = str(tmp_path / name)
    subprocess.run([exe, *map(str, args), prefix], check=True, timeout=120)
    return (np.fromfile(prefix + ".u0.bin", dtype=np.float64),
            np.fromfile(prefix + ".u.bin", dtype=np.float64))


def boundary_nodes(c, face, N, cells):
    """(element, dof) of a face's boundary nodes in the per-node state order:
    (element position along the face) * (N+1) + (node position)."""
    n, (nx, ny) = N + 1, cells
    out = []
    for idx in range(ny if face < 2 else nx):
        for q in range(n):
            if face < 2:
                out.append(((0 if face == 0 else nx - 1) * ny + idx, q * n + (0 if face == 0 else n - 1)))
            else:
                out.append((idx * ny + (0 if face == 2 else ny - 1), (0 if face == 2 else n - 1) * n + q))
    return out


def double_mach_state(x, y, t, gamma=1.4):
    if x < 1.0 / 6.0 + (y + 20.0 * t) / math.sqrt(3.0):
        rho, vx, vy, p = 8.0, 8.25 * math.cos(math.pi / 6.0), -8.25 * math.sin(math.pi / 6.0), 116.5
    else:
        rho, vx, vy, p = 1.4, 0.0, 0.0, 1.0
    return [rho, rho * vx, rho * vy, p / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy)]


@pytest.mark.gpu
def test_cpp_double_mach_reflection_matches_oracle(gpu, tmp_path):
    """Time-dependent Dirichlet inflow (left, top), outflow (right) and the
    mixed Dirichlet/slip-wall bottom, composed as in the reference's example;
    the oracle is stepped with the boundary states re-evaluated at each
    SSPRK3 stage time (t, t + dt, t + dt/2, SSPRK3.hpp:159-181)."""
    N, nx, ny, steps = 4, 24, 6, 20
    dt = 0.2 * min(4.0 / nx, 1.0 / ny) / ((2 * (N + 1) - 1) * 12.0)
    print(f"case cpp dmr {nx}x{ny} steps={steps}", flush=True)
    u0, got = run_example("double_mach_reflection", tmp_path, nx, ny, steps, dt)
    c = O.Case(2, N, (nx, ny), [0.0, 0.0], [4.0, 1.0], periodic=[0, 0], volume=O.SHOCK_CAPTURING,
               alpha_max=0.5, alpha_min=0.001)
    x = c.coordinates()
    nodes = {f: boundary_nodes(c, f, N, (nx, ny)) for f in (0, 2, 3)}
    c.set_boundary(1, O.BC_OUTFLOW)

    def set_time(t):
        for f in (0, 3):
            c.set_boundary(f, O.BC_DIRICHLET,
                           values=np.array([double_mach_state(*x[e, d], t) for e, d in nodes[f]]))
        c.set_boundary(2, O.BC_MIXED_DIRICHLET_SLIP_WALL, values=np.array(
            [double_mach_state(*x[e, d], t) + [float(x[e, d, 0] < 1.0 / 6.0)] for e, d in nodes[2]]))

    u = u0.reshape(c.shape).copy()
    t = 0.0
    for _ in range(steps):
        set_time(t)
        u1 = 1.0 * u + dt * c.rhs(u)
        set_time(t + dt)
        u2 = 3.0 / 4.0 * u + 1.0 / 4.0 * u1 + 1.0 / 4.0 * dt * c.rhs(u1)
        set_time(t + 0.5 * dt)
        u = 1.0 / 3.0 * u + 2.0 / 3.0 * u2 + 2.0 / 3.0 * dt * c.rhs(u2)
        t += dt
    got = got.reshape(u.shape)
    assert np.isfinite(got).all()
    assert np.linalg.norm(got - u) / np.linalg.norm(u) <= 1e-10


def astro_jet_state(x, y, t, gamma=1.4):
    if t > 0.0 and abs(x + 0.5) < 1e-8 and abs(y) < 0.05:
        rho, vx = 5.0, 800.0
    else:
        rho, vx = 0.5, 0.0
    p = 0.4127
    return [rho, rho * vx, 0.0, p / (gamma - 1.0) + 0.5 * rho * vx * vx]


@pytest.mark.gpu
def test_cpp_astro_jet_matches_oracle(gpu, tmp_path):
    """The jet inflow (examples/Euler2D/astro_jet): Dirichlet faces left and
    right whose left state switches to the rho = 5, vx = 800 jet on
    |y| < 0.05 for t > 0, periodic in y; the oracle re-evaluates the states
    at each SSPRK3 stage time, so the first stage of the first step sees the
    ambient state and every later one the jet."""
    N, nx, ny, steps, dt = 3, 16, 16, 20, 1e-8
    print(f"case cpp astro jet {nx}x{ny} steps={steps}", flush=True)
    u0, got = run_example("astro_jet", tmp_path, nx, ny, steps, dt)
    c = O.Case(2, N, (nx, ny), [-0.5, -0.5], [0.5, 0.5], periodic=[0, 1], volume=O.SHOCK_CAPTURING,
               alpha_max=0.3, alpha_min=1e-4)
    x = c.coordinates()
    nodes = {f: boundary_nodes(c, f, N, (nx, ny)) for f in (0, 1)}

    def set_time(t):
        for f in (0, 1):
            c.set_boundary(f, O.BC_DIRICHLET,
                           values=np.array([astro_jet_state(*x[e, d], t) for e, d in nodes[f]]))

    u = u0.reshape(c.shape).copy()
    t = 0.0
    for _ in range(steps):
        set_time(t)
        u1 = 1.0 * u + dt * c.rhs(u)
        set_time(t + dt)
        u2 = 3.0 / 4.0 * u + 1.0 / 4.0 * u1 + 1.0 / 4.0 * dt * c.rhs(u1)
        set_time(t + 0.5 * dt)
        u = 1.0 / 3.0 * u + 2.0 / 3.0 * u2 + 2.0 / 3.0 * dt * c.rhs(u2)
        t += dt
    got = got.reshape(u.shape)
    assert np.isfinite(got).all()
    # the jet has entered: the left column's momentum is no longer zero
    assert np.abs(u[..., 1]).max() > 1.0
    assert np.linalg.norm(got - u) / np.linalg.norm(u) <= 1e-10


@pytest.mark.gpu
def test_cpp_riemann2d_matches_oracle(gpu, tmp_path):
    """The four-quadrant Riemann problem (examples/Euler2D/riemann2d): outflow
    on every face, N = 5, the example's CFL-0.2 time step."""
    N, nx, ny, steps = 5, 20, 20, 30
    dt = 0.2 * min(1.0 / nx, 1.0 / ny) / ((2 * (N + 1) - 1) * 2.0)
    print(f"case cpp riemann2d {nx}x{ny} steps={steps}", flush=True)
    u0, got = run_example("riemann2d", tmp_path, nx, ny, steps)
    c = O.Case(2, N, (nx, ny), [0.0, 0.0], [1.0, 1.0], periodic=[0, 0], volume=O.SHOCK_CAPTURING,
               alpha_max=0.5, alpha_min=0.001)
    for f in range(4):
        c.set_boundary(f, O.BC_OUTFLOW)
    _, alpha = c.rhs(u0.reshape(c.shape), with_alpha=True)
    assert 0 < int((alpha > 0).sum()) < c.E
    ref = c.ssprk3(u0.reshape(c.shape), dt, steps)
    got = got.reshape(ref.shape)
    assert np.isfinite(got).all()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-10


@pytest.mark.gpu
def test_cpp_viscous_shock_tube_matches_oracle(gpu, tmp_path):
    """Navier-Stokes adiabatic walls (no slip on three sides, slip on top)
    through the reference-style interface."""
    N, nx, ny, steps = 3, 16, 8, 30
    dt = 0.4 * min(1.0 / nx, 0.5 / ny) / ((2 * N + 1) * 7.0)
    print(f"case cpp viscous shock tube {nx}x{ny} steps={steps}", flush=True)
    u0, got = run_example("viscous_shock_tube", tmp_path, nx, ny, steps, dt)
    c = O.Case(2, N, (nx, ny), [0.0, 0.0], [1.0, 0.5], periodic=[0, 0], eq=O.NAVIER_STOKES,
               mu=1e-3, prandtl=0.73, volume=O.SHOCK_CAPTURING, alpha_max=1.0, alpha_min=0.1)
    for f, k in enumerate([O.BC_NS_NOSLIP_WALL] * 3 + [O.BC_NS_SLIP_WALL]):
        c.set_boundary(f, k)
    ref = c.ssprk3(u0.reshape(c.shape), dt, steps)
    got = got.reshape(ref.shape)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-10


@pytest.mark.gpu
def test_cpp_observers_example(gpu, tmp_path):
    """The reference-style driver with observers (taylor_green_observers):
    PrintObserver lines every 5 steps, VTU files at steps 0, 10 and 20, the
    NaN check and the volume-average CSV before every step and at the end,
    whose kinetic energies match the Python path's on-device analysis of the
    same run."""
    import torch

    import dgsem_amd as dg
    from dgsem_amd import tgv

    exe = os.path.join(ROOT, "dgsem_amd", "examples", "taylor_green_observers")
    assert os.path.exists(exe), "build the library first (dgsem_amd/build.py)"
    cells, N, steps = 4, 3, 20
    prefix = str(tmp_path / "tgv")
    print("case cpp observers tgv 4^3 N=3 20 steps", flush=True)
    r = subprocess.run([exe, str(cells), str(N), str(steps), prefix], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert [l.split()[1] for l in lines if l.startswith("Iter:")] == ["0", "5", "10", "15"]
    assert any(l.startswith(f"Finished at iter {steps}") for l in lines)
    dt = float([l for l in lines if l.startswith("dt ")][0].split()[1])
    for k in (0, 10, 20):
        assert os.path.getsize(f"{prefix}_step{k}.vtu") > 0
    rows = open(prefix + "_averages.csv").read().splitlines()
    assert rows[0] == "step,time,volume_average_kinetic_energy,volume_average_entropy"
    data = np.array([[float(x) for x in row.split(",")] for row in rows[1:]])
    assert data.shape == (steps + 1, 4)  # before every step, and once at the end
    # the same run through the Python interface
    s = dg.StructuredSolver(dg.SolverConfig(dim=3, polydeg=N, cells=(cells,) * 3,
                                            domain_left=(0, 0, 0),
                                            domain_right=(2 * math.pi,) * 3))
    u = s.from_reference_layout(tgv.tgv_state(s.node_coordinates()))
    rk = dg.SSPRK3(s)
    ke = []
    for n in range(steps + 1):
        ke.append(s.analyze(u)["mean_kinetic_energy"])
        if n < steps:
            rk.step(u, dt)
    torch.cuda.synchronize()
    assert np.allclose(data[:, 2], ke, rtol=1e-13, atol=0.0)
    assert np.all(np.diff(data[:, 0]) == 1) and np.isfinite(data).all()
