"""Differentiable slice loop: Ez of one step as a torch.autograd.Function of the beam
macro-particle weights, backed by the adjoint slice loop of the C ABI
(hpq_adjoint_forward / hpq_adjoint_backward: checkpointed forward pass, reverse march of
hand-written vector-Jacobian products). No counterpart in the reference, whose slice
loop has no derivative.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .simulation import Simulation


class SliceLoopEz(torch.autograd.Function):
    """Ez[slice, j, i] over the valid box for every slice of one step, as a function of the
    beam weights w (one per macro-particle, in the beam's slice-sorted order)."""

    @staticmethod
    def forward(ctx, beam_w: torch.Tensor, sim: Simulation, checkpoint_every: int = 32,
                mg_tol: float = 1e-10):
        sim.beam["w"][: beam_w.numel()].copy_(beam_w.detach().to(sim.beam["w"].dtype))
        ez = sim.capture("Ez")
        sim.begin_step()
        sim.adjoint_forward(checkpoint_every)
        torch.cuda.synchronize(sim.device)
        out = ez.clone()
        sim.clear_captures()
        ctx.sim = sim
        ctx.mg_tol = mg_tol
        ctx.n = beam_w.numel()
        return out

    @staticmethod
    def backward(ctx, grad_ez: torch.Tensor):
        sim = ctx.sim
        w_adj = torch.zeros(max(ctx.n, 1), dtype=torch.float64, device=sim.device)
        sim.adjoint_backward(grad_ez.detach().to(torch.float64).contiguous(), w_adj, ctx.mg_tol)
        return w_adj[: ctx.n], None, None, None


def slice_loop_ez(beam_w: torch.Tensor, sim: Simulation, checkpoint_every: int = 32,
                  mg_tol: float = 1e-10) -> torch.Tensor:
    return SliceLoopEz.apply(beam_w, sim, checkpoint_every, mg_tol)


def knot_profile(z: torch.Tensor, knots: torch.Tensor, z0: float, z1: float) -> torch.Tensor:
    """Piecewise-linear current profile through len(knots) equally spaced knots on [z0, z1],
    constant beyond the ends: the factor applied to the weights of beam particles at z."""
    n = knots.numel()
    u = ((z - z0) / (z1 - z0) * (n - 1)).clamp(0, n - 1)
    i0 = u.floor().long().clamp(0, max(n - 2, 0))
    f = u - i0.to(u.dtype)
    if n == 1:
        return knots[0].expand_as(z)
    return knots[i0] * (1 - f) + knots[i0 + 1] * f


def on_axis_ez_variance(ez: torch.Tensor, slices: slice) -> torch.Tensor:
    """Variance of Ez on the axis cell (ny//2, nx//2) over the given slices: the flatness
    of the accelerating field seen by a witness beam."""
    ny, nx = ez.shape[1], ez.shape[2]
    v = ez[slices, ny // 2, nx // 2]
    return ((v - v.mean()) ** 2).mean()
