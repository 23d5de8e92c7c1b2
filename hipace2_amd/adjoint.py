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


ADJOINT_FIELDS = ("Ez", "jz_beam")


class SliceLoop(torch.autograd.Function):
    """Fields of one step on every slice (the valid box, [slice, j, i]) as a function of the
    beam weights w (one per macro-particle, in the beam's slice-sorted order): Ez, and
    jz_beam when asked for, so that a loss may depend on both (the beam's energy gain,
    sum Ez * jz_beam dx dy dz, is one).

    The checkpoints of the forward pass live in the Simulation's context, and the backward pass
    recomputes from them with the beam weights in the Simulation: a backward is valid only for
    the last forward run on that Simulation, with its weights unchanged. Each forward takes a
    new generation number; backward raises if another forward ran since, or if the weights
    changed."""

    @staticmethod
    def forward(ctx, beam_w: torch.Tensor, sim: Simulation, checkpoint_every: int = 32,
                mg_tol: float = 1e-10, fields: tuple = ("Ez",)):
        for f in fields:
            if f not in ADJOINT_FIELDS:
                raise ValueError(f"slice_loop: no adjoint for field {f!r} (have {ADJOINT_FIELDS})")
        w = beam_w.detach().to(sim.beam["w"].dtype)
        sim.beam["w"][: beam_w.numel()].copy_(w)
        caps = {f: sim.capture(f) for f in fields}
        sim.begin_step()
        sim.adjoint_forward(checkpoint_every)
        torch.cuda.synchronize(sim.device)
        outs = tuple(caps[f].clone() for f in fields)
        sim.clear_captures()
        ctx.set_materialize_grads(False)
        ctx.sim = sim
        ctx.mg_tol = mg_tol
        ctx.n = beam_w.numel()
        ctx.fields = tuple(fields)
        ctx.generation = sim.adjoint_generation
        ctx.w = w.clone()
        return outs

    @staticmethod
    def backward(ctx, *grads):
        sim = ctx.sim
        if sim.adjoint_generation != ctx.generation:
            raise RuntimeError("slice_loop: another forward pass ran on this Simulation since; its "
                               "checkpoints replaced this one's")
        if not torch.equal(sim.beam["w"][: ctx.n], ctx.w):
            raise RuntimeError("slice_loop: the beam weights changed between the forward and the backward pass")
        seeds = {f: (g.detach().to(torch.float64).contiguous() if g is not None else None)
                 for f, g in zip(ctx.fields, grads)}
        w_adj = torch.zeros(max(ctx.n, 1), dtype=torch.float64, device=sim.device)
        sim.adjoint_backward(seeds.get("Ez"), w_adj, ctx.mg_tol, jz_beam_seed=seeds.get("jz_beam"))
        return w_adj[: ctx.n], None, None, None, None


def slice_loop(beam_w: torch.Tensor, sim: Simulation, checkpoint_every: int = 32, mg_tol: float = 1e-10,
               fields: tuple = ("Ez", "jz_beam")):
    """The captured fields (a tuple in the order of `fields`) of one step, differentiable in
    the beam weights."""
    return SliceLoop.apply(beam_w, sim, checkpoint_every, mg_tol, tuple(fields))


def slice_loop_ez(beam_w: torch.Tensor, sim: Simulation, checkpoint_every: int = 32,
                  mg_tol: float = 1e-10) -> torch.Tensor:
    return SliceLoop.apply(beam_w, sim, checkpoint_every, mg_tol, ("Ez",))[0]


def energy_gain(ez: torch.Tensor, jz_beam: torch.Tensor, slices: slice, cfg) -> torch.Tensor:
    """Energy the beam gains per unit time over the given slices: the sum over those slices of
    the in-situ quantity [Ez*jz_beam] = sum over valid cells of Ez * jz_beam * dx dy dz
    (Fields::InSituComputeDiags, src/fields/Fields.cpp:1320-1346). jz_beam = q w v_z S(x) S(y)
    is negative for electrons moving forward, so a decelerating field (Ez > 0) gives a
    positive product: the quantity is the power the beam delivers to the wake, and the
    witness gains energy when it is negative."""
    return -(ez[slices] * jz_beam[slices]).sum() * (cfg.dx * cfg.dy * cfg.dz)


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


def witness_beam(cfg, num_particles: int, z0: float, z1: float, sigma_r: float, n_peak: float,
                 uz: float, seed: int = 7):
    """Equal-weight witness macro-particles, uniform in zeta on [z0, z1] and Gaussian in x, y,
    of peak density n_peak, sorted into zeta-slices like the drive (weight as
    make_fixed_weight_beam: density times volume over dx dy dz N)."""
    from .inputs import BeamArrays
    rng = np.random.Generator(np.random.Philox(seed))
    n = int(num_particles)
    z = rng.uniform(z0, z1, n)
    x = rng.normal(0.0, sigma_r, n)
    y = rng.normal(0.0, sigma_r, n)
    total = n_peak * (sigma_r * np.sqrt(2.0 * np.pi)) ** 2 * (z1 - z0) / (cfg.dx * cfg.dy * cfg.dz)
    w = np.full(n, total / n)
    box = np.floor((z - cfg.zlo) / cfg.dz).astype(np.int64)
    inside = (box >= 0) & (box < cfg.nz)
    order = np.argsort(np.where(inside, box, cfg.nz), kind="stable")
    order = order[inside[order]]
    off = np.zeros(cfg.nz + 1, np.int64)
    off[1:] = np.cumsum(np.bincount(box[inside], minlength=cfg.nz))
    zeros = np.zeros(order.size)
    return BeamArrays(x[order], y[order], z[order], zeros, zeros.copy(), np.full(order.size, uz), w[order],
                      np.ones(order.size, np.int32), off)


class BeamLoading:
    """Drive + witness beam whose witness current profile is piecewise linear through knots:
    the witness weights are w0 * knot_profile(zeta; knots). Maps knots to the weights of the
    merged, slice-sorted beam (the drive's weights stay fixed)."""

    def __init__(self, cfg, drive, witness, z0: float, z1: float):
        from .inputs import merge_beams
        self.beam = merge_beams(cfg, [drive, witness])
        self.z0, self.z1 = z0, z1
        idx = []
        for s in range(cfg.nz):
            start = int(self.beam.off[s]) + int(drive.off[s + 1] - drive.off[s])
            idx.append(np.arange(start, start + int(witness.off[s + 1] - witness.off[s])))
        self.witness_index = np.concatenate(idx) if idx else np.zeros(0, np.int64)
        self.witness_z = witness.z
        self.witness_w0 = witness.w
        didx = [np.arange(int(self.beam.off[s]), int(self.beam.off[s]) + int(drive.off[s + 1] - drive.off[s]))
                for s in range(cfg.nz)]
        self.drive_index = np.concatenate(didx) if didx else np.zeros(0, np.int64)
        self.drive_z = drive.z
        self.drive_w0 = drive.w
        self.base_w = self.beam.w.copy()
        self.base_w[self.witness_index] = 0.0
        self.s0 = int(np.floor((z0 - cfg.zlo) / cfg.dz))
        self.s1 = int(np.floor((z1 - cfg.zlo) / cfg.dz))

    def weights(self, knots: torch.Tensor, drive_knots: Optional[torch.Tensor] = None,
                drive_range: tuple = (-4.0, 4.0)) -> torch.Tensor:
        """Weights of the merged beam: the witness's are w0 * knot_profile(zeta; knots) on
        [z0, z1]; with drive_knots, the drive's current profile is shaped the same way,
        w0 * knot_profile(zeta; drive_knots) on drive_range (constant beyond its ends)."""
        dev = knots.device
        base = torch.as_tensor(self.base_w, device=dev)
        zw = torch.as_tensor(self.witness_z, device=dev)
        w0 = torch.as_tensor(self.witness_w0, device=dev)
        idx = torch.as_tensor(self.witness_index, device=dev)
        w = base.index_put((idx,), w0 * knot_profile(zw, knots, self.z0, self.z1))
        if drive_knots is not None:
            zd = torch.as_tensor(self.drive_z, device=dev)
            d0 = torch.as_tensor(self.drive_w0, device=dev)
            didx = torch.as_tensor(self.drive_index, device=dev)
            w = w.index_put((didx,), d0 * knot_profile(zd, drive_knots, drive_range[0], drive_range[1]))
        return w

    def witness_slices(self) -> slice:
        return slice(max(self.s0, 0), self.s1 + 1)
