"""Differentiation through the HIP right-hand side and SSPRK3 steps.

The derivatives are the hand-written tangent and adjoint kernels of
libdgsem_amd.so (dgsem_rhs_jvp / dgsem_rhs_vjp / dgsem_ssprk3_step_vjp,
include/dgsem_amd.h); this module only orchestrates them:

* RHS: torch.autograd.Function for L(u) (forward dgsem_rhs, backward the
  adjoint kernel);
* kinetic_energy_objective: J(u0) = sum_{n < nsteps} dt KE(u_n), the
  time-integrated kinetic energy sampled before each step, with its gradient
  from a reverse sweep over checkpoints stored every `every` steps (the steps
  between two checkpoints are recomputed, their stage states kept for the
  sweep);
* KineticEnergyObjective: the same as a torch.autograd.Function of u0.
"""
from __future__ import annotations

import math

import torch

from .solver import StructuredSolver


class RHS(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, solver: StructuredSolver):
        u = u.detach().contiguous()
        du = solver.empty()
        solver.calc_rhs(u, du)
        ctx.solver = solver
        ctx.save_for_backward(u)
        return du

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        return ctx.solver.rhs_vjp(u, g.contiguous()), None


def _step(solver, u, dt, u1, u2, out):
    solver.ssprk3_stage(1, dt, None, u, u1)
    solver.ssprk3_stage(2, dt, u, u1, u2)
    solver.ssprk3_stage(3, dt, u, u2, out)


def kinetic_energy_trajectory(solver: StructuredSolver, u0, dt: float, nsteps: int, every: int = 0):
    """Runs nsteps SSPRK3 steps from u0.  Returns (J, checkpoints) where
    checkpoints maps step index -> state for every multiple of `every`
    (0 = only the initial state)."""
    ke = torch.empty(max(nsteps, 1), dtype=torch.float64, device="cuda")
    ckpt = {0: u0.detach().clone()}
    u, u1, u2, nxt = u0.detach().clone(), solver.empty(), solver.empty(), solver.empty()
    for n in range(nsteps):
        solver.kinetic_energy(u, ke[n:n + 1])
        _step(solver, u, dt, u1, u2, nxt)
        u, nxt = nxt, u
        if every and (n + 1) % every == 0 and n + 1 < nsteps:
            ckpt[n + 1] = u.clone()
    vals = ke[:nsteps].cpu().tolist()
    return dt * math.fsum(vals), ckpt


def kinetic_energy_objective(solver: StructuredSolver, u0, dt: float, nsteps: int,
                             every: int = 20):
    """(J, dJ/du0) with checkpoints every `every` steps."""
    J, ckpt = kinetic_energy_trajectory(solver, u0, dt, nsteps, every)
    return J, kinetic_energy_gradient(solver, ckpt, dt, nsteps, every)


def kinetic_energy_gradient(solver: StructuredSolver, ckpt: dict, dt: float, nsteps: int,
                            every: int):
    lam = torch.zeros(solver.shape, dtype=torch.float64, device="cuda")
    t1, t2 = solver.empty(), solver.empty()
    every = every or nsteps
    starts = list(range(0, nsteps, every))
    for s0 in reversed(starts):
        s1 = min(s0 + every, nsteps)
        # recompute the segment, keeping the stage states of every step
        states = []
        u = ckpt[s0].clone()
        for _ in range(s0, s1):
            u1, u2 = solver.empty(), solver.empty()
            solver.ssprk3_stage(1, dt, None, u, u1)
            solver.ssprk3_stage(2, dt, u, u1, u2)
            states.append((u, u1, u2))
            nxt = solver.empty()
            solver.ssprk3_stage(3, dt, u, u2, nxt)
            u = nxt
        for n in range(s1 - 1, s0 - 1, -1):
            un, u1, u2 = states[n - s0]
            # lambda_n = (du_{n+1}/du_n)^T lambda_{n+1} + dt dKE(u_n)/du_n
            solver.step_vjp(dt, un, u1, u2, lam, out=lam, tmp1=t1, tmp2=t2)
            solver.kinetic_energy_grad(un, dt, lam)
        del states
    return lam


class KineticEnergyObjective(torch.autograd.Function):
    """J(u0) = sum_{n < nsteps} dt KE(u_n) as a differentiable function of u0."""

    @staticmethod
    def forward(ctx, u0, solver: StructuredSolver, dt: float, nsteps: int, every: int = 20):
        J, ckpt = kinetic_energy_trajectory(solver, u0.detach().contiguous(), dt, nsteps, every)
        ctx.solver, ctx.dt, ctx.nsteps, ctx.every, ctx.ckpt = solver, dt, nsteps, every, ckpt
        return torch.tensor(J, dtype=torch.float64, device=u0.device)

    @staticmethod
    def backward(ctx, g):
        lam = kinetic_energy_gradient(ctx.solver, ctx.ckpt, ctx.dt, ctx.nsteps, ctx.every)
        ctx.ckpt = None
        return g * lam, None, None, None, None


def conservative_from_velocity(rho, v, p, gamma: float = 1.4):
    """u = (rho, rho v, p/(gamma-1) + rho |v|^2 / 2) in the internal layout;
    rho, p: [E][nn], v: [E][dim][nn].  Plain torch ops, so autograd carries
    dJ/du0 to dJ/dv."""
    m = rho[:, None, :] * v
    E = p / (gamma - 1.0) + 0.5 * rho * (v * v).sum(dim=1)
    return torch.cat([rho[:, None, :], m, E[:, None, :]], dim=1).contiguous()
