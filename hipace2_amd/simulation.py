"""Host driver of the slice loop over the C ABI.

`Simulation` owns the device buffers of one run (slice store, plasma SoA, beam arrays,
diagnostics), allocated through torch on the GPU; every computation is a call into
libhpq.so on torch's current HIP stream. It mirrors the reference's Hipace object for
this path: InitData (plasma and beam), Evolve's per-step reset and neutralising
background, SolveOneSlice over the ζ-slices from head to tail.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Dict, Optional

import numpy as np
import torch

from . import _abi
from .inputs import BC_CODES, BeamArrays, SliceConfig, lattice_density

PLASMA_REALS = ["x", "y", "w", "ux", "uy", "psi", "x_prev", "y_prev", "ux_half_step", "uy_half_step",
                "psi_half_step"]


def make_params(cfg: SliceConfig) -> _abi.Params:
    (plx, ply), (phx, phy) = cfg.bounds
    p = _abi.Params()
    p.nx, p.ny, p.nz = cfg.nx, cfg.ny, cfg.nz
    p.xlo, p.xhi, p.ylo, p.yhi, p.zlo, p.zhi = cfg.xlo, cfg.xhi, cfg.ylo, cfg.yhi, cfg.zlo, cfg.zhi
    p.depos_order, p.derivative_type = cfg.depos_order, cfg.derivative_type
    p.explicit_solver, p.deposit_rho = int(cfg.explicit_solver), int(cfg.deposit_rho)
    p.do_beam_jx_jy = int(cfg.do_beam_jx_jy)
    p.neutralize = int(cfg.neutralize and cfg.has_plasma)
    p.mg_tol_rel, p.mg_tol_abs, p.mg_maxiter = cfg.mg_tol_rel, cfg.mg_tol_abs, cfg.mg_maxiter
    p.particle_bc = BC_CODES[cfg.particle_bc]
    p.plo[0], p.plo[1], p.phi[0], p.phi[1] = plx, ply, phx, phy
    p.pc_tol, p.pc_maxiter, p.pc_mix = cfg.pc_tol, cfg.pc_maxiter, cfg.pc_mix
    p.plasma_charge, p.plasma_mass, p.max_qsa = cfg.plasma_charge, cfg.plasma_mass, cfg.max_qsa
    p.bin_period = cfg.bin_period
    p.precision = _abi.PRECISION[cfg.precision]
    p.poisson_solver = _abi.POISSON_SOLVER[cfg.poisson_solver]
    p.salame = int(cfg.salame)
    p.deposition = _abi.DEPOSITION[cfg.deposition]
    p.tile_size = cfg.tile_size
    p.reorder_period = cfg.reorder_period
    p.mixed_sums = _abi.MIXED_SUMS[cfg.mixed_sums]
    p.mg_guess = _abi.MG_GUESS[cfg.mg_guess]
    p.mg_guess_iters = cfg.mg_guess_iters
    return p


def make_profile(cfg: SliceConfig) -> _abi.PlasmaProfile:
    pr = _abi.PlasmaProfile()
    pr.ppc[0], pr.ppc[1] = cfg.ppc
    pr.density, pr.ramp_a = cfg.density, cfg.ramp_a
    pr.radius = cfg.plasma_radius if math.isfinite(cfg.plasma_radius) else 1e300
    for k in range(3):
        pr.u_mean[k] = cfg.u_mean[k]
    return pr


class Simulation:
    def __init__(self, cfg: SliceConfig, beam: Optional[BeamArrays] = None, device: str = "cuda:0"):
        if not torch.cuda.is_available():
            raise RuntimeError("hipace2_amd needs a GPU (HIP device)")
        self.cfg = cfg
        self.device = torch.device(device)
        self.L = _abi.lib()
        self._params = make_params(cfg)
        h = C.c_void_p()
        rc = self.L.hpq_create(C.byref(self._params), C.byref(h))
        self.ctx = h
        _abi.check(self.ctx, rc)
        self.set_stream(torch.cuda.current_stream(self.device))
        arr = (C.c_int32 * len(_abi.LAYOUT_NAMES))()
        self.L.hpq_layout(self.ctx, arr)
        self.layout = dict(zip(_abi.LAYOUT_NAMES, list(arr)))
        self.ncomp = self.layout["ncomp"]
        self.ldx, self.ldy = cfg.nx + 4, cfg.ny + 4
        self.store = torch.zeros((self.ncomp, self.ldy, self.ldx), dtype=torch.float64, device=self.device)
        # plasma SoA
        self._profile = make_profile(cfg)
        if cfg.has_plasma and cfg.density_expr is not None:
            # <plasma>.density(x,y,z) evaluated at every lattice position
            self._density_table = torch.as_tensor(lattice_density(cfg)).to(self.device)
            self._profile.density_table = self._density_table.data_ptr()
        n = int(self.L.hpq_plasma_count(self.ctx, C.byref(self._profile))) if cfg.has_plasma else 0
        self.np = n
        self.plasma = {k: torch.zeros(max(n, 1), dtype=torch.float64, device=self.device) for k in PLASMA_REALS}
        self.plasma["valid"] = torch.zeros(max(n, 1), dtype=torch.int32, device=self.device)
        self._pdesc = _abi.Plasma(*[self.plasma[k].data_ptr() for k in PLASMA_REALS + ["valid"]], n)
        self.checksum_buf = torch.zeros(self.ncomp, dtype=torch.float64, device=self.device)
        _abi.check(self.ctx, self.L.hpq_set_checksums(self.ctx, C.c_void_p(self.checksum_buf.data_ptr())))
        self._captures: Dict[str, torch.Tensor] = {}
        self.set_beam(beam if beam is not None else BeamArrays.empty(cfg.nz))

    def __del__(self):
        try:
            torch.cuda.synchronize(self.device)
            self.L.hpq_destroy(self.ctx)
        except Exception:
            pass

    # ---- setup ----------------------------------------------------------------------
    def set_stream(self, stream):
        self.stream = stream
        _abi.check(self.ctx, self.L.hpq_set_stream(self.ctx, C.c_void_p(stream.cuda_stream)))

    def set_beam(self, b: BeamArrays):
        self.beam_host = b
        dev = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(self.device)
        n = b.n
        self.beam = {k: dev(getattr(b, k) if n else np.zeros(1)) for k in ("x", "y", "z", "ux", "uy", "uz", "w")}
        self.beam["valid"] = dev(b.valid if n else np.zeros(1, np.int32), torch.int32)
        if n and b.salame is not None:
            self.beam["salame"] = dev(b.salame, torch.int32)
        self._boff = (C.c_int64 * (self.cfg.nz + 1))(*[int(v) for v in b.off])
        self._bdesc = _abi.Beam(*[self.beam[k].data_ptr() for k in ("x", "y", "z", "ux", "uy", "uz", "w", "valid")],
                                C.cast(self._boff, C.POINTER(C.c_int64)), n, b.charge, b.mass)

    def _beam_state(self):
        if "nsub" not in self.beam:
            self.beam["nsub"] = torch.zeros(max(self.beam_host.n, 1), dtype=torch.int32, device=self.device)
        self._bsdesc = _abi.BeamState(*[self.beam[k].data_ptr() for k in
                                        ("x", "y", "z", "ux", "uy", "uz", "w", "valid", "nsub")],
                                      C.cast(self._boff, C.POINTER(C.c_int64)), self.beam_host.n,
                                      self.beam_host.charge, self.beam_host.mass)
        ext = self.beam_host.external
        if ext is not None:
            for k, v in enumerate(np.asarray(ext, dtype=np.float64).reshape(24)):
                self._bsdesc.external_fields[k] = float(v)
            self._bsdesc.use_external_fields = 1
        return self._bsdesc

    def beam_begin_step(self):
        """Reset the beam sub-cycle counters and the slipped-particle list (hpq_beam_begin_step)."""
        _abi.check(self.ctx, self.L.hpq_beam_begin_step(self.ctx, C.byref(self._beam_state())))

    def advance_beam_slice(self, islice: int, dt: float, n_subcycles: int = 10, do_z_push: bool = True):
        """AdvanceBeamParticlesSlice + shiftSlippedParticles for one slice (hpq_advance_beam_slice)."""
        _abi.check(self.ctx, self.L.hpq_advance_beam_slice(self.ctx, C.byref(self._beam_state()), islice,
                                                           C.c_void_p(self.store.data_ptr()), float(dt),
                                                           int(n_subcycles), int(do_z_push)))

    def sort_beam(self):
        """Re-sort the beam by zeta-slice on the device (hpq_sort_beam_slices); updates the
        slice offsets used by deposition and push."""
        _abi.check(self.ctx, self.L.hpq_sort_beam_slices(self.ctx, C.byref(self._beam_state()), self._boff))
        self._beam_sorted = True   # the per-particle SALAME flags no longer follow the particles
        self.beam_host.off = np.array(list(self._boff), dtype=np.int64)
        self._bdesc.off = C.cast(self._boff, C.POINTER(C.c_int64))

    def run_step_with_beam(self, dt: float, n_subcycles: int = 10, do_z_push: bool = True,
                           init_plasma: bool = True):
        """One time step with the beam pushed after every slice (hpq_run_step_with_beam)."""
        self.begin_step(init_plasma)
        _abi.check(self.ctx, self.L.hpq_run_step_with_beam(self.ctx, C.byref(self._pdesc),
                                                           C.byref(self._beam_state()),
                                                           C.c_void_p(self.store.data_ptr()), float(dt),
                                                           int(n_subcycles), int(do_z_push)))

    def run_steps(self, nsteps: int, dt: float, n_subcycles: int = 10, do_z_push: bool = True):
        """Several time steps: beam re-sorted by slice between steps, plasma re-injected each
        step (src/Hipace.cpp:426-480); with a SALAME beam, SALAME runs in the first step."""
        for step in range(nsteps):
            if step > 0:
                self.sort_beam()
            if self.cfg.salame and step == 0:
                self.enable_salame()
            self.run_step_with_beam(dt, n_subcycles, do_z_push)
            if self.cfg.salame and step == 0:
                self.disable_salame()

    # ---- SALAME -------------------------------------------------------------------------
    def enable_salame(self, n_iter: Optional[int] = None, do_advance: Optional[bool] = None,
                      relative_tolerance: Optional[float] = None, target=None):
        """Run SALAME on the following slices (hpq_set_salame) with the deck's settings unless
        given: the weights of the SALAME beam's particles (<beam>.do_salame) are rescaled on the
        device, slice by slice, to flatten Ez. `target(zeta, zeta_initial, Ez_initial)` replaces
        hipace.salame_Ez_target; the default keeps Ez_initial. The last weight factor of every
        slice is in self.salame_weight_factor afterwards."""
        from .inputs import parse_function
        cfg = self.cfg
        if "salame" not in self.beam:
            raise ValueError("SALAME needs beam particles of a beam with do_salame")
        if getattr(self, "_beam_sorted", False):
            raise ValueError("SALAME runs in the first time step, before the beam is re-sorted by slice")
        if target is None and cfg.salame_target.replace(" ", "") != "Ez_initial":
            target = parse_function(cfg.salame_target, ["zeta", "zeta_initial", "Ez_initial"])
        d = _abi.Salame()
        d.mask = self.beam["salame"].data_ptr()
        d.w = self.beam["w"].data_ptr()
        d.valid = self.beam["valid"].data_ptr()
        d.n_iter = cfg.salame_n_iter if n_iter is None else int(n_iter)
        d.do_advance = int(cfg.salame_do_advance if do_advance is None else do_advance)
        d.relative_tolerance = cfg.salame_relative_tolerance if relative_tolerance is None else relative_tolerance
        self._salame_cb = None
        if target is not None:
            self._salame_cb = _abi.SALAME_TARGET(lambda z, z0, e0, _u: float(target(z, z0, e0)))
            d.target = self._salame_cb
        self.salame_weight_factor = np.zeros(cfg.nz)
        d.weight_factor = self.salame_weight_factor.ctypes.data_as(C.POINTER(C.c_double))
        _abi.check(self.ctx, self.L.hpq_set_salame(self.ctx, C.byref(d)))
        self._salame_desc = d

    def disable_salame(self):
        _abi.check(self.ctx, self.L.hpq_set_salame(self.ctx, None))
        self._salame_desc = None

    def init_plasma(self):
        if self.np:
            _abi.check(self.ctx, self.L.hpq_init_plasma(self.ctx, C.byref(self._profile), C.byref(self._pdesc)))

    def set_plasma(self, d: Dict[str, np.ndarray]):
        for k in PLASMA_REALS + ["valid"]:
            src = torch.as_tensor(np.ascontiguousarray(d[k]))
            self.plasma[k][: self.np].copy_(src.to(self.plasma[k].dtype))

    def capture(self, name: str) -> torch.Tensor:
        """Record the valid region of component `name` at every slice's diagnostics point."""
        buf = torch.zeros((self.cfg.nz, self.cfg.ny, self.cfg.nx), dtype=torch.float64, device=self.device)
        self._captures[name] = buf
        _abi.check(self.ctx, self.L.hpq_add_capture(self.ctx, self.layout[name], C.c_void_p(buf.data_ptr())))
        return buf

    def clear_captures(self):
        """Stop recording every captured component (hpq_clear_captures)."""
        _abi.check(self.ctx, self.L.hpq_clear_captures(self.ctx))
        self._captures.clear()

    # ---- slice loop -------------------------------------------------------------------
    def begin_step(self, init_plasma: bool = True):
        if init_plasma:
            self.init_plasma()
        self.checksum_buf.zero_()
        _abi.check(self.ctx, self.L.hpq_begin_step(self.ctx, C.byref(self._pdesc),
                                                   C.c_void_p(self.store.data_ptr())))

    def solve_slice(self, islice: int):
        _abi.check(self.ctx, self.L.hpq_solve_one_slice(self.ctx, C.byref(self._pdesc), C.byref(self._bdesc),
                                                        C.c_void_p(self.store.data_ptr()), islice))

    def run_slices(self, first: Optional[int] = None, last: int = 0):
        first = self.cfg.nz - 1 if first is None else first
        _abi.check(self.ctx, self.L.hpq_run_slices(self.ctx, C.byref(self._pdesc), C.byref(self._bdesc),
                                                   C.c_void_p(self.store.data_ptr()), first, last))

    def run_step(self, first: Optional[int] = None, last: int = 0):
        self.begin_step()
        self.run_slices(first, last)

    # ---- results ----------------------------------------------------------------------
    def checksums(self) -> Dict[str, float]:
        torch.cuda.synchronize(self.device)
        v = self.checksum_buf.cpu().numpy()
        return {k: float(v[i]) for k, i in self.layout.items() if k != "ncomp" and i >= 0}

    def stats(self) -> Dict[str, float]:
        out = (C.c_double * 4)()
        _abi.check(self.ctx, self.L.hpq_stats(self.ctx, out))
        return dict(n_qsa=out[0], mg_iters=out[1], pc_iters=out[2], pushes=out[3])

    def trace(self, kinds, on: bool = True):
        """Enable HIP-event timing of the named kernel families (_abi.TRACE)."""
        for k in kinds:
            _abi.check(self.ctx, self.L.hpq_trace_enable(self.ctx, _abi.TRACE[k], int(on)))

    def trace_reset(self):
        _abi.check(self.ctx, self.L.hpq_trace_reset(self.ctx))

    def trace_read(self, kind: str):
        """(total kernel milliseconds, launches) since the last reset."""
        ms = C.c_double(0)
        n = C.c_int64(0)
        _abi.check(self.ctx, self.L.hpq_trace_read(self.ctx, _abi.TRACE[kind], C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ---- in-situ reduced diagnostics -------------------------------------------------------
    def enable_insitu(self, fields: bool = True, plasma: bool = True, beam: bool = True,
                      plasma_radius: float = math.inf, beam_radius: float = math.inf):
        """Per-slice reductions on the device during the slice loop (hpq_set_insitu): the
        fields after the field solves, the plasma at the start of each slice, the beam
        particles of the slice before its push. plasmas/beams.insitu_radius cut the particles."""
        nz = self.cfg.nz
        f64 = dict(dtype=torch.float64, device=self.device)
        i32 = dict(dtype=torch.int32, device=self.device)
        buf: Dict[str, torch.Tensor] = {}
        d = _abi.Insitu()
        if fields:
            buf["field_rdata"] = torch.zeros((_abi.INSITU_FIELD_NR, nz), **f64)
            buf["field_sum"] = torch.zeros(_abi.INSITU_FIELD_NR, **f64)
        for kind, nr, radius, on in (("plasma", _abi.INSITU_PLASMA_NR, plasma_radius, plasma),
                                     ("beam", _abi.INSITU_BEAM_NR, beam_radius, beam)):
            if on:
                buf[kind + "_rdata"] = torch.zeros((nr, nz), **f64)
                buf[kind + "_idata"] = torch.zeros(nz, **i32)
                buf[kind + "_sum"] = torch.zeros(nr, **f64)
                buf[kind + "_isum"] = torch.zeros(1, **i32)
                setattr(d, kind + "_radius", float(radius))
        for k, t in buf.items():
            setattr(d, k, t.data_ptr())
        _abi.check(self.ctx, self.L.hpq_set_insitu(self.ctx, C.byref(d)))
        self._insitu, self._insitu_desc = buf, d

    def disable_insitu(self):
        _abi.check(self.ctx, self.L.hpq_set_insitu(self.ctx, None))
        self._insitu = {}

    def insitu_arrays(self) -> Dict[str, np.ndarray]:
        """Host copies of the in-situ arrays: <group>_rdata (NR, nz), _idata (nz), _sum (NR), _isum (1)."""
        torch.cuda.synchronize(self.device)
        return {k: v.cpu().numpy() for k, v in getattr(self, "_insitu", {}).items()}

    def reset_insitu(self):
        for v in getattr(self, "_insitu", {}).values():
            v.zero_()

    def write_insitu(self, step: int, time: float, field_prefix: str = "diags/field_insitu",
                     plasma_prefix: str = "diags/plasma_insitu", beam_prefix: str = "diags/insitu",
                     plasma_name: str = "plasma", beam_name: str = "beam", rank: int = 0):
        """Append this step's records (Fields/Plasma/BeamParticleContainer::InSituWriteToFile)
        and clear the arrays, as the reference does after writing."""
        import os
        from . import insitu as io
        a = self.insitu_arrays()
        cfg = self.cfg
        dxdydz = cfg.dx * cfg.dy * cfg.dz
        if "field_rdata" in a:
            os.makedirs(field_prefix, exist_ok=True)
            io.append_record(io.file_name(field_prefix, "fields", rank),
                             io.field_record(time, step, cfg.zlo, cfg.zhi, a["field_rdata"], a["field_sum"]))
        for kind, prefix, name, charge, mass in (
                ("plasma", plasma_prefix, plasma_name, cfg.plasma_charge, cfg.plasma_mass),
                ("beam", beam_prefix, beam_name, self.beam_host.charge, self.beam_host.mass)):
            if kind + "_rdata" in a:
                os.makedirs(prefix, exist_ok=True)
                io.append_record(io.file_name(prefix, name, rank),
                                 io.species_record(kind, time, step, charge, mass, cfg.zlo, cfg.zhi, dxdydz,
                                                   a[kind + "_rdata"], a[kind + "_idata"], a[kind + "_sum"],
                                                   int(a[kind + "_isum"][0])))
        self.reset_insitu()

    def plasma_host(self) -> Dict[str, np.ndarray]:
        torch.cuda.synchronize(self.device)
        return {k: v[: self.np].cpu().numpy() for k, v in self.plasma.items()}

    def store_host(self) -> np.ndarray:
        torch.cuda.synchronize(self.device)
        return self.store.cpu().numpy()

    # ---- kernel-level entry points (the reference's free functions) -----------------------
    def deposit_current(self, comps, ions: bool = False) -> int:
        arr = (C.c_int32 * 6)(*[self.layout.get(c, -1) if isinstance(c, str) else c for c in comps])
        nq = C.c_int64(0)
        _abi.check(self.ctx, self.L.hpq_deposit_current(self.ctx, C.byref(self._pdesc),
                                                        C.c_void_p(self.store.data_ptr()), arr, int(ions),
                                                        C.byref(nq)))
        return nq.value

    def explicit_deposition(self):
        _abi.check(self.ctx, self.L.hpq_explicit_deposition(self.ctx, C.byref(self._pdesc),
                                                            C.c_void_p(self.store.data_ptr())))

    def deposit_current_slice(self, islice: int, cjx, cjy, cjz):
        f = lambda c: self.layout.get(c, -1) if isinstance(c, str) else c
        arr = (C.c_int32 * 3)(f(cjx), f(cjy), f(cjz))
        _abi.check(self.ctx, self.L.hpq_deposit_current_slice(self.ctx, C.byref(self._bdesc), islice,
                                                              C.c_void_p(self.store.data_ptr()), arr))

    # ---- adjoint building blocks (vector-Jacobian products) ------------------------------
    def new_plasma_adj(self) -> Dict[str, torch.Tensor]:
        """Zeroed adjoint arrays (ADJ_NAMES) for the current plasma."""
        return {k: torch.zeros(self.np, dtype=torch.float64, device=self.device) for k in _abi.ADJ_NAMES}

    def new_store_adj(self) -> torch.Tensor:
        """Zeroed adjoint planes in the slice-store layout."""
        return torch.zeros_like(self.store)

    @staticmethod
    def _adj_desc(a: Dict[str, torch.Tensor]):
        return _abi.PlasmaAdj(*[a[k].data_ptr() for k in _abi.ADJ_NAMES])

    def deposit_current_vjp(self, comps, store_adj: torch.Tensor, adj: Dict[str, torch.Tensor]):
        f = lambda c: self.layout.get(c, -1) if isinstance(c, str) else c
        arr = (C.c_int32 * 6)(*[f(c) for c in comps])
        d = self._adj_desc(adj)
        _abi.check(self.ctx, self.L.hpq_deposit_current_vjp(self.ctx, C.byref(self._pdesc),
                                                            C.c_void_p(store_adj.data_ptr()), arr, C.byref(d)))

    def explicit_deposition_vjp(self, store_adj: torch.Tensor, adj: Dict[str, torch.Tensor]):
        d = self._adj_desc(adj)
        _abi.check(self.ctx, self.L.hpq_explicit_deposition_vjp(self.ctx, C.byref(self._pdesc),
                                                                C.c_void_p(self.store.data_ptr()),
                                                                C.c_void_p(store_adj.data_ptr()), C.byref(d)))

    def advance_plasma_particles_vjp(self, out_adj: Dict[str, torch.Tensor], in_adj: Dict[str, torch.Tensor],
                                     store_adj: torch.Tensor):
        o, i = self._adj_desc(out_adj), self._adj_desc(in_adj)
        _abi.check(self.ctx, self.L.hpq_advance_plasma_particles_vjp(
            self.ctx, C.byref(self._pdesc), C.c_void_p(self.store.data_ptr()), C.byref(o), C.byref(i),
            C.c_void_p(store_adj.data_ptr())))

    adjoint_generation = 0

    def adjoint_forward(self, checkpoint_every: int = 32):
        """One step's slices with checkpoints for the backward march (after begin_step); every
        call starts a new generation of checkpoints (adjoint_generation)."""
        self.adjoint_generation += 1
        _abi.check(self.ctx, self.L.hpq_adjoint_forward(self.ctx, C.byref(self._pdesc), C.byref(self._bdesc),
                                                        C.c_void_p(self.store.data_ptr()), int(checkpoint_every)))

    def adjoint_backward(self, ez_seed: Optional[torch.Tensor], beam_w_adj: torch.Tensor, mg_tol: float = 1e-10,
                         jz_beam_seed: Optional[torch.Tensor] = None):
        """Adds d loss / d (beam weights) to beam_w_adj given d loss / d Ez and d loss / d jz_beam
        ([slice][j][i] each; None: zero)."""
        seeds = []
        for t in (ez_seed, jz_beam_seed):
            if t is not None:
                assert t.is_contiguous() and t.dtype == torch.float64
                assert t.numel() == self.cfg.nz * self.cfg.ny * self.cfg.nx
            seeds.append(C.c_void_p(t.data_ptr()) if t is not None else None)
        _abi.check(self.ctx, self.L.hpq_adjoint_backward(self.ctx, C.byref(self._pdesc), C.byref(self._bdesc),
                                                         C.c_void_p(self.store.data_ptr()), seeds[0], seeds[1],
                                                         C.c_void_p(beam_w_adj.data_ptr()), float(mg_tol)))

    def advance_plasma_particles(self, temp_slice: bool = False):
        _abi.check(self.ctx, self.L.hpq_advance_plasma_particles(self.ctx, C.byref(self._pdesc),
                                                                 C.c_void_p(self.store.data_ptr()),
                                                                 int(temp_slice)))

    def solve_poisson_psi_exmby_eypbx_ez_bz(self):
        _abi.check(self.ctx, self.L.hpq_solve_poisson_psi_exmby_eypbx_ez_bz(self.ctx,
                                                                            C.c_void_p(self.store.data_ptr())))

    def solve_poisson_bxby(self):
        _abi.check(self.ctx, self.L.hpq_solve_poisson_bxby(self.ctx, C.c_void_p(self.store.data_ptr())))

    def initialize_sxsy_with_beam(self):
        _abi.check(self.ctx, self.L.hpq_initialize_sxsy_with_beam(self.ctx, C.c_void_p(self.store.data_ptr())))

    def mg_solve1(self, tol_rel=1e-4, tol_abs=2.2250738585072014e-308, maxiter=200):
        plane = self.ldx * self.ldy
        base = self.store.data_ptr()
        it = C.c_int(0)
        rc = self.L.hpq_mg_solve1(self.ctx, C.c_void_p(base + 8 * plane * self.layout["Bx"]),
                                  C.c_void_p(base + 8 * plane * self.layout["Sy"]),
                                  C.c_void_p(base + 8 * plane * self.layout["chi"]), plane, tol_rel, tol_abs,
                                  maxiter, C.byref(it))
        return rc, it.value

    def bin_tiles(self, tile: int):
        """hpq_bin_tiles on the current plasma: (perm, offsets) as int32 device tensors."""
        n = self.np
        ntx = (self.cfg.nx + 4 + tile - 1) // tile
        nty = (self.cfg.ny + 4 + tile - 1) // tile
        perm = torch.empty(max(n, 1), dtype=torch.int32, device=self.device)
        offsets = torch.empty(ntx * nty + 2, dtype=torch.int32, device=self.device)
        _abi.check(self.ctx, self.L.hpq_bin_tiles(self.ctx, C.byref(self._pdesc), tile,
                                                  C.c_void_p(perm.data_ptr()), C.c_void_p(offsets.data_ptr())))
        return perm[:n], offsets

    def bin_tiles_noalloc(self, tile: int, perm: torch.Tensor, offsets: torch.Tensor):
        """hpq_bin_tiles_noalloc on the current plasma into caller-allocated int32 device
        tensors (perm: at least max(n, 1) entries; offsets: ntile_x * ntile_y + 2)."""
        assert perm.dtype == torch.int32 and offsets.dtype == torch.int32
        assert perm.numel() >= max(self.np, 1)
        _abi.check(self.ctx, self.L.hpq_bin_tiles_noalloc(self.ctx, C.byref(self._pdesc), tile,
                                                          C.c_void_p(perm.data_ptr()),
                                                          C.c_void_p(offsets.data_ptr())))

    def poisson_dirichlet(self, staging: torch.Tensor, lhs: torch.Tensor):
        assert staging.is_contiguous() and staging.numel() == self.cfg.nx * self.cfg.ny
        assert lhs.is_contiguous() and lhs.numel() == self.ldx * self.ldy
        _abi.check(self.ctx, self.L.hpq_poisson_dirichlet(self.ctx, C.c_void_p(staging.data_ptr()),
                                                          C.c_void_p(lhs.data_ptr())))
