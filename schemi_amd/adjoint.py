"""Gradients through the GPU step (no counterpart in the reference).

* ``Steps`` -- a ``torch.autograd.Function`` over the C ABI: forward runs n
  SSP-RK2 steps of fixed size on the device and keeps a checkpoint of the
  state every ``every`` steps in device memory; backward recomputes each
  segment from its checkpoint and applies ``schemi_gpu_adjoint_step`` in
  reverse order (flux branches frozen at the forward states, dt fixed).
* ``MixingWidth`` -- J = sum_i V_i Y_i (1 - Y_i), Y = rho_species / rho, and
  its gradient, from ``schemi_gpu_mixing_width``.
* ``rt_modes_state`` -- the 2-D Rayleigh-Taylor initial state (heavy over
  light, hydrostatic pressure) as a differentiable torch function of the
  amplitudes a_k of the interface modes cos(2 pi k x / Lx), k = 1..K.

Composed, ``MixingWidth(Steps(rt_modes_state(a)))`` gives dJ/da by
``torch.autograd``: the initial-state Jacobian from torch, the step adjoint
from the HIP kernels.  Tensors are FP64 on the GPU in the state layout
(Ns + 5 rows of ncells values, cells x fastest).
"""
from __future__ import annotations

import math

import torch

from .case import Case
from .cases import cv_of
from .enums import boundaryConditionType as BCT
from .enums import flowSolver, typeOfTVDLimiter
from .gpu import GpuSolver, periodic

ADJ_STEP = 0
ADJ_STAGE = 1


def _ptr(t: torch.Tensor) -> int:
    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
    return t.data_ptr()


class Steps(torch.autograd.Function):
    """U_N = S^N(U_0) for a solver created with a constant time step dt."""

    @staticmethod
    def forward(ctx, U0: torch.Tensor, solver: GpuSolver, nsteps: int, dt: float,
                every: int = 20):
        U0 = U0.detach().contiguous()
        torch.cuda.synchronize()
        solver.state_device(src=_ptr(U0))
        cps = []
        done = 0
        while done < nsteps:
            cp = torch.empty_like(U0)
            solver.state_device(dst=_ptr(cp))
            cps.append(cp)
            k = min(every, nsteps - done)
            solver.run(k, dt)
            done += k
        out = torch.empty_like(U0)
        solver.state_device(dst=_ptr(out))
        ctx.solver, ctx.nsteps, ctx.dt, ctx.every = solver, nsteps, dt, every
        ctx.cps = cps
        return out

    @staticmethod
    def backward(ctx, gout: torch.Tensor):
        solver, dt, every = ctx.solver, ctx.dt, ctx.every
        lam = gout.detach().to(torch.float64).contiguous().clone()
        torch.cuda.synchronize()
        nseg = len(ctx.cps)
        for s in range(nseg - 1, -1, -1):
            k = min(every, ctx.nsteps - s * every)
            # forward states of the segment, recomputed from its checkpoint
            states = [ctx.cps[s]]
            solver.state_device(src=_ptr(ctx.cps[s]))
            for _ in range(k - 1):
                solver.run(1, dt)
                st = torch.empty_like(lam)
                solver.state_device(dst=_ptr(st))
                states.append(st)
            for t in range(k - 1, -1, -1):
                solver.state_device(src=_ptr(states[t]))
                solver.adjoint_step(dt, _ptr(lam), _ptr(lam), ADJ_STEP)
        ctx.cps = None
        return lam, None, None, None, None


class MixingWidth(torch.autograd.Function):
    """J(U) = sum_i V_i Y_i (1 - Y_i) with Y = rho_species / rho."""

    @staticmethod
    def forward(ctx, U: torch.Tensor, solver: GpuSolver, species: int = 0):
        U = U.detach().contiguous()
        grad = torch.empty_like(U)
        torch.cuda.synchronize()
        solver.state_device(src=_ptr(U))
        J = solver.mixing_width(species, _ptr(grad))
        ctx.save_for_backward(grad)
        return torch.tensor(J, dtype=torch.float64, device=U.device)

    @staticmethod
    def backward(ctx, gJ: torch.Tensor):
        (grad,) = ctx.saved_tensors
        return gJ * grad, None, None


def rt_case(nx: int = 256, ny: int = 1024, periodic_x: bool = True) -> Case:
    """2-D Rayleigh-Taylor set-up of the gradient workload: heavy (M 87) over
    light (M 29), both gamma 1.4, on [0,0.25]x[0,1] (Lz = 1/ny), gravity -1
    in y, HLLC + van Leer (exactly first order in 2-D), slip walls in y,
    periodic (or slip) in x."""
    case = Case(dim=2, N=(nx, ny, 1), L=(0.25, 1.0, 1.0 / ny), M=[87.0, 29.0],
                Cv=[cv_of(1.4), cv_of(1.4)], flux=flowSolver.HLLC,
                limiter=typeOfTVDLimiter.vanLeer, gravity=True, g=(0.0, -1.0, 0.0),
                courant=0.4, min_time=1e-15)
    for pt in (3, 4):
        case.set_bc(pt, BCT.slip)
    if periodic_x:
        periodic(case, 0)
    else:
        for pt in (0, 1):
            case.set_bc(pt, BCT.slip)
    return case


def rt_modes_state(case: Case, a: torch.Tensor, width_cells: float = 2.0) -> torch.Tensor:
    """Conserved state (Ns + 5, n) of the Rayleigh-Taylor set-up with interface
    eta(x) = 0.5 + sum_k a_k cos(2 pi k x / Lx): heavy mass fraction
    Y = (1 + tanh((y - eta) / w)) / 2 with w = width_cells / ny, density
    1 / (Y / 3 + (1 - Y) / 1), pressure hydrostatic per column
    (dp/dy = -rho, trapezoidal between cell centres, p = 2.5 at y = 0.5), at
    rest; E = p / (R / Cv) (both species gamma 1.4).  Differentiable in a."""
    nx, ny = case.N[0], case.N[1]
    dev, f64 = a.device, torch.float64
    Lx = case.L[0]
    dy = case.L[1] / ny
    xs = (torch.arange(nx, dtype=f64, device=dev) + 0.5) * (Lx / nx)
    ys = (torch.arange(ny, dtype=f64, device=dev) + 0.5) * dy
    k = torch.arange(1, a.shape[0] + 1, dtype=f64, device=dev)
    eta = 0.5 + (a[:, None] * torch.cos(2.0 * math.pi * k[:, None] * xs[None, :] / Lx)).sum(0)
    w = width_cells * dy
    Y = 0.5 * (1.0 + torch.tanh((ys[:, None] - eta[None, :]) / w))  # (ny, nx)
    rho = 1.0 / (Y / 3.0 + (1.0 - Y) / 1.0)
    # hydrostatic pressure from y = 0.5 outwards
    j0 = int(math.floor(0.5 / dy))
    p0 = 2.5 - rho[j0] * (ys[j0] - 0.5)
    up = 0.5 * (rho[j0:-1] + rho[j0 + 1:]) * dy  # p[j] - p[j+1], j >= j0
    dn = 0.5 * (rho[:j0] + rho[1:j0 + 1]) * dy   # p[j] - p[j+1], j < j0
    p_up = p0[None, :] - torch.cumsum(up, 0)
    p_dn = p0[None, :] + torch.flip(torch.cumsum(torch.flip(dn, [0]), 0), [0])
    p = torch.cat([p_dn, p0[None, :], p_up], 0)
    rA = Y * rho
    rB = (1.0 - Y) * rho
    cvm = (rA / case.M[0] * case.Cv[0] + rB / case.M[1] * case.Cv[1]) / \
        (rA / case.M[0] + rB / case.M[1])
    E = p / (case.R / cvm)
    z = torch.zeros_like(rho)
    return torch.stack([rA, rB, rA + rB, z, z, z, E]).reshape(7, nx * ny).contiguous()


def mixing_width_gradient(case: Case, a, nsteps: int, dt: float, every: int = 20,
                          precision: int = 0, species: int = 0):
    """J(a) and dJ/da for the Rayleigh-Taylor workload: the forward in
    ``precision`` (its states are the linearisation points), the adjoint in
    FP64.  Returns (J, grad) as float and numpy array."""
    solver = GpuSolver(case, precision, const_dt=dt)
    try:
        at = torch.as_tensor(a, dtype=torch.float64, device="cuda").clone().requires_grad_(True)
        U0 = rt_modes_state(case, at)
        UN = Steps.apply(U0, solver, nsteps, dt, every)
        J = MixingWidth.apply(UN, solver, species)
        (g,) = torch.autograd.grad(J, at)
        return float(J.detach()), g.detach().cpu().numpy()
    finally:
        solver.close()


def mixing_width_forward(case: Case, a, nsteps: int, dt: float, precision: int = 0,
                         species: int = 0) -> float:
    """J(a) by a plain forward run (for finite differences)."""
    solver = GpuSolver(case, precision, const_dt=dt)
    try:
        at = torch.as_tensor(a, dtype=torch.float64, device="cuda")
        U0 = rt_modes_state(case, at).contiguous()
        torch.cuda.synchronize()
        solver.state_device(src=_ptr(U0))
        solver.run(nsteps, dt)
        return solver.mixing_width(species)
    finally:
        solver.close()
