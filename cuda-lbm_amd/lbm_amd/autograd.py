"""PyTorch-ROCm autograd bridge over the discrete adjoint of liblbm_hip.so.

`simulate(force, u0)` runs T fused steps from f = f_eq(rho0, u0) under the body-force field
`force` and returns the velocity snapshot u_T; its backward runs the hand-written adjoint
kernels (reverse streaming + transposed BGK/MRT Jacobian, recomputing each segment from
an FP32 checkpoint taken every K steps) and returns dL/dforce and dL/du0. Tensors stay on
the GPU; the library runs on PyTorch's current stream.
"""
from __future__ import annotations

import torch

from .solver import Config, Solver


def _run_forward(sim, force, u0):
    """Upload the inputs and run the checkpointing forward on PyTorch's current stream."""
    s = sim.solver
    s.set_stream(torch.cuda.current_stream().cuda_stream)
    s.set_force_field(force.detach().contiguous())
    s.init_from_macros(sim.rho0, u0.detach().contiguous())
    s.adjoint_forward(sim.nsteps, sim.checkpoint_every)
    sim._run += 1
    return sim._run


class _LbmRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, force, u0, sim):
        run = _run_forward(sim, force, u0)
        uT = torch.empty_like(u0)
        sim.solver.get_macros_device(u=uT)
        ctx.sim = sim
        ctx.run = run
        ctx.save_for_backward(force, u0)
        return uT

    @staticmethod
    def backward(ctx, grad_u):
        sim = ctx.sim
        if sim._run != ctx.run:
            # The solver's checkpoints and force field belong to a later call of `sim` (for
            # example a validation forward): run this node's forward again from its own inputs.
            force, u0 = ctx.saved_tensors
            ctx.run = _run_forward(sim, force, u0)
        s = sim.solver
        s.set_stream(torch.cuda.current_stream().cuda_stream)
        gF = torch.empty_like(grad_u)
        gU = torch.empty_like(grad_u)
        s.adjoint_backward(grad_u.contiguous(), gF, gU)
        return gF, gU, None


class Simulation:
    """A differentiable run of `nsteps` D2Q9 steps (BGK or MRT, fully periodic, FP32)."""

    def __init__(self, cfg: Config, nsteps: int, checkpoint_every: int = 100, rho0=None):
        self.cfg = cfg
        self.nsteps = int(nsteps)
        self.checkpoint_every = int(checkpoint_every)
        self.solver = Solver(cfg)
        self._run = 0  # forward runs so far; the solver holds the checkpoints of the last one
        dev = torch.device("cuda", torch.cuda.current_device())
        self.rho0 = (torch.ones(cfg.ny, cfg.nx, dtype=torch.float32, device=dev) if rho0 is None
                     else torch.as_tensor(rho0, dtype=torch.float32, device=dev).contiguous())

    def __call__(self, force: torch.Tensor, u0: torch.Tensor) -> torch.Tensor:
        return _LbmRun.apply(force, u0, self)

    def close(self):
        self.solver.close()


def recover_force(sim: Simulation, u_target: torch.Tensor, u0: torch.Tensor, iters: int = 200, lr: float = 2e-6,
                  force0: torch.Tensor | None = None):
    """Fit a body-force field so that the run reproduces `u_target` (Adam on the L2 misfit).
    Returns (force, losses)."""
    force = torch.zeros_like(u_target) if force0 is None else force0.clone()
    force.requires_grad_(True)
    opt = torch.optim.Adam([force], lr=lr)
    losses = []
    for _ in range(iters):
        opt.zero_grad()
        uT = sim(force, u0)
        loss = torch.sum((uT - u_target) ** 2)
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    return force.detach(), losses
