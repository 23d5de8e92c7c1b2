"""ctypes binding of the oracle (CPU FP64 restatement of the reference slice loop).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")

LAYOUT_NAMES = ["ncomp", "next_jx_beam", "next_jy_beam", "chi", "Sy", "Sx", "ExmBy", "EypBx", "Ez", "Bx",
                "By", "Bz", "Psi", "jx_beam", "jy_beam", "jz_beam", "jx", "jy", "rhomjz", "rho",
                "prev_jx_beam", "prev_jy_beam", "ions_rhomjz", "next_jx", "next_jy", "jz", "prev_Bx",
                "prev_By", "prev_jx", "prev_jy", "pciter_Bx", "pciter_By",
                "salame_Ez_target", "salame_Ez_no_salame", "salame_Ez", "salame_jx", "salame_jy",
                "salame_jz_beam", "salame_Bx", "salame_By", "salame_Sy", "salame_Sx", "salame_Sy_back",
                "salame_Sx_back"]

PLASMA_REALS = ["x", "y", "w", "ux", "uy", "psi", "x_prev", "y_prev", "ux_half_step", "uy_half_step",
                "psi_half_step"]


class OrcConfig(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xlo", C.c_double), ("xhi", C.c_double), ("ylo", C.c_double), ("yhi", C.c_double),
                ("zlo", C.c_double), ("zhi", C.c_double),
                ("depos_order", C.c_int), ("derivative_type", C.c_int), ("explicit_solver", C.c_int),
                ("deposit_rho", C.c_int), ("do_beam_jx_jy", C.c_int),
                ("mg_tol_rel", C.c_double), ("mg_tol_abs", C.c_double), ("mg_maxiter", C.c_int),
                ("particle_bc", C.c_int), ("plo", C.c_double * 2), ("phi", C.c_double * 2),
                ("pc_tol", C.c_double), ("pc_maxiter", C.c_int), ("pc_mix", C.c_double),
                ("has_plasma", C.c_int), ("ppc", C.c_int * 2),
                ("density", C.c_double), ("ramp_a", C.c_double), ("plasma_radius", C.c_double),
                ("u_mean", C.c_double * 3),
                ("plasma_charge", C.c_double), ("plasma_mass", C.c_double), ("max_qsa", C.c_double),
                ("neutralize", C.c_int), ("nthreads", C.c_int), ("salame", C.c_int), ("tile_size", C.c_int)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)
    return LIB_PATH


_lib = None
_lib_path = LIB_PATH


def use_library(path: str) -> None:
    """Load the oracle from another build of the same sources (e.g. one compiled for the host
    it runs on); must come before the first call into the oracle."""
    global _lib_path
    if _lib is not None and path != _lib_path:
        raise RuntimeError("oracle library already loaded")
    _lib_path = path


def lib():
    global _lib
    if _lib is None:
        if _lib_path == LIB_PATH and not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(_lib_path)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.orc_sim_create.restype = vp
        L.orc_sim_create.argtypes = [C.POINTER(OrcConfig)]
        L.orc_sim_destroy.argtypes = [vp]
        L.orc_sim_set_density_table.argtypes = [vp, dp]
        L.orc_sim_ncomp.argtypes = [vp]
        L.orc_sim_layout.argtypes = [vp, ip]
        L.orc_sim_set_beam.argtypes = [vp, C.c_int64] + [dp] * 7 + [C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                                             C.c_double, C.c_double]
        L.orc_sim_add_capture.argtypes = [vp, C.c_int, dp]
        L.orc_sim_clear_captures.argtypes = [vp]
        L.orc_sim_begin_step.argtypes = [vp]
        L.orc_sim_solve_slice.argtypes = [vp, C.c_int]
        L.orc_sim_run_step.argtypes = [vp, C.c_int, C.c_int]
        L.orc_sim_run_step.restype = C.c_double
        L.orc_sim_checksums.argtypes = [vp, dp]
        L.orc_sim_stats.argtypes = [vp, dp]
        L.orc_sim_get_store.argtypes = [vp, dp]
        L.orc_sim_set_store.argtypes = [vp, dp]
        L.orc_sim_plasma_size.argtypes = [vp]
        L.orc_sim_plasma_size.restype = C.c_int64
        L.orc_sim_get_plasma.argtypes = [vp, dp, C.POINTER(C.c_int32)]
        L.orc_sim_set_plasma.argtypes = [vp, C.c_int64, dp, C.POINTER(C.c_int32)]
        L.orc_deposit_plasma.argtypes = [vp, ip, C.c_int, C.POINTER(C.c_int64)]
        L.orc_explicit_deposit.argtypes = [vp]
        L.orc_deposit_beam.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_push_plasma.argtypes = [vp, C.c_int]
        i32p = C.POINTER(C.c_int32)
        L.orc_push_beam_slice.argtypes = [vp] + [dp] * 7 + [i32p, i32p, C.c_double, C.c_double, C.c_long, C.c_long,
                                                          i32p, C.c_long, i32p, C.c_long, C.c_int, C.c_double, C.c_int,
                                                          C.c_int, dp]
        L.orc_sim_set_beam_external.argtypes = [vp, dp]
        L.orc_sim_set_salame.argtypes = [vp, C.POINTER(C.c_int32), C.c_int, C.c_int, C.c_double, SALAME_TARGET,
                                         vp, dp]
        L.orc_push_beam_slice.restype = C.c_long
        L.orc_sim_run_step_with_beam.argtypes = [vp, C.c_double, C.c_int, C.c_int]
        L.orc_sim_sort_beam.argtypes = [vp]
        L.orc_sim_get_beam.argtypes = [vp] + [dp] * 7 + [i32p, C.POINTER(C.c_int64)]
        L.orc_solve_poisson_psi_ez_bz.argtypes = [vp]
        L.orc_solve_poisson_bxby.argtypes = [vp]
        L.orc_initialize_sxsy_with_beam.argtypes = [vp]
        L.orc_mg_solve1_store.argtypes = [vp, C.c_double, C.c_double, C.c_int, ip]
        L.orc_dst1.argtypes = [dp, dp, C.c_int]
        L.orc_poisson.argtypes = [C.POINTER(OrcConfig), dp, dp]
        L.orc_shape.argtypes = [C.c_int, C.c_double, C.c_int, dp]
        L.orc_dshape.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


SALAME_TARGET = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p)


def make_config(cfg, nthreads: int = 0) -> OrcConfig:
    from hipace2_amd.inputs import BC_CODES
    (plx, ply), (phx, phy) = cfg.bounds
    c = OrcConfig()
    c.nx, c.ny, c.nz = cfg.nx, cfg.ny, cfg.nz
    c.xlo, c.xhi, c.ylo, c.yhi, c.zlo, c.zhi = cfg.xlo, cfg.xhi, cfg.ylo, cfg.yhi, cfg.zlo, cfg.zhi
    c.depos_order, c.derivative_type = cfg.depos_order, cfg.derivative_type
    c.explicit_solver, c.deposit_rho = int(cfg.explicit_solver), int(cfg.deposit_rho)
    c.do_beam_jx_jy = int(cfg.do_beam_jx_jy)
    c.mg_tol_rel, c.mg_tol_abs, c.mg_maxiter = cfg.mg_tol_rel, cfg.mg_tol_abs, cfg.mg_maxiter
    c.particle_bc = BC_CODES[cfg.particle_bc]
    c.plo[0], c.plo[1], c.phi[0], c.phi[1] = plx, ply, phx, phy
    c.pc_tol, c.pc_maxiter, c.pc_mix = cfg.pc_tol, cfg.pc_maxiter, cfg.pc_mix
    c.has_plasma = int(cfg.has_plasma)
    c.ppc[0], c.ppc[1] = cfg.ppc
    c.density, c.ramp_a, c.plasma_radius = cfg.density, cfg.ramp_a, cfg.plasma_radius
    for k in range(3):
        c.u_mean[k] = cfg.u_mean[k]
    c.plasma_charge, c.plasma_mass, c.max_qsa = cfg.plasma_charge, cfg.plasma_mass, cfg.max_qsa
    c.neutralize = int(cfg.neutralize)
    c.nthreads = nthreads
    c.salame = int(getattr(cfg, "salame", False))
    c.tile_size = int(getattr(cfg, "tile_size", 32))
    return c


class OracleSim:
    """One slice-loop simulation on the CPU (the checker)."""

    def __init__(self, cfg, beam=None, nthreads: int = 0, density_table=None):
        """density_table: <plasma>.density(x,y,z) at every lattice position (the layout of
        hpq_plasma_profile.density_table), required when cfg.density_expr is set."""
        self.cfg = cfg
        self._c = make_config(cfg, nthreads)
        self.h = lib().orc_sim_create(C.byref(self._c))
        arr = (C.c_int * len(LAYOUT_NAMES))()
        lib().orc_sim_layout(self.h, arr)
        self.layout = dict(zip(LAYOUT_NAMES, list(arr)))
        self.ncomp = self.layout["ncomp"]
        self.ng = 2
        self.ldx, self.ldy = cfg.nx + 4, cfg.ny + 4
        self._captures = {}
        self._beam = None
        if beam is not None:
            self.set_beam(beam)
        self._density_table = None
        if getattr(cfg, "density_expr", None) is not None and density_table is None:
            raise ValueError("a density expression needs its lattice table")
        if density_table is not None:
            self._density_table = np.ascontiguousarray(density_table, dtype=np.float64)
            if self._density_table.size != cfg.nx * cfg.ny * cfg.ppc[0] * cfg.ppc[1]:
                raise ValueError("density table size")
            lib().orc_sim_set_density_table(self.h, _dp(self._density_table))

    def __del__(self):
        try:
            lib().orc_sim_destroy(self.h)
        except Exception:
            pass

    def set_beam(self, b):
        self._beam = [np.ascontiguousarray(a, dtype=np.float64) for a in (b.x, b.y, b.z, b.ux, b.uy, b.uz, b.w)]
        self._bvalid = np.ascontiguousarray(b.valid, dtype=np.int32)
        self._boff = np.ascontiguousarray(b.off, dtype=np.int64)
        lib().orc_sim_set_beam(self.h, b.n, *[_dp(a) for a in self._beam],
                               self._bvalid.ctypes.data_as(C.POINTER(C.c_int32)),
                               self._boff.ctypes.data_as(C.POINTER(C.c_int64)), b.charge, b.mass)
        ext = getattr(b, "external", None)
        self._bext = None if ext is None else np.ascontiguousarray(ext, dtype=np.float64).reshape(24)
        lib().orc_sim_set_beam_external(self.h, None if self._bext is None else _dp(self._bext))

    def enable_salame(self, mask, n_iter=5, do_advance=True, relative_tolerance=1e-4, target=None):
        """SalameModule on the following slices for the beam particles with mask 1; the last
        weight factor of every slice goes to self.salame_weight_factor."""
        self._sal_mask = np.ascontiguousarray(mask, dtype=np.int32)
        self._sal_cb = SALAME_TARGET(lambda z, z0, e0, _u: float(target(z, z0, e0))) if target else SALAME_TARGET()
        self.salame_weight_factor = np.zeros(self.cfg.nz)
        lib().orc_sim_set_salame(self.h, self._sal_mask.ctypes.data_as(C.POINTER(C.c_int32)), int(n_iter),
                                 int(do_advance), float(relative_tolerance), self._sal_cb, None,
                                 _dp(self.salame_weight_factor))

    def disable_salame(self):
        lib().orc_sim_set_salame(self.h, None, 0, 0, 0.0, SALAME_TARGET(), None, None)

    def capture(self, name: str) -> np.ndarray:
        """Record the valid region of component `name` at the diagnostics point of every slice."""
        buf = np.zeros((self.cfg.nz, self.cfg.ny, self.cfg.nx))
        self._captures[name] = buf
        lib().orc_sim_add_capture(self.h, self.layout[name], _dp(buf))
        return buf

    def clear_captures(self):
        """Stop recording every captured component."""
        lib().orc_sim_clear_captures(self.h)
        self._captures.clear()

    def run_step(self, first=None, last=0) -> float:
        first = self.cfg.nz - 1 if first is None else first
        return lib().orc_sim_run_step(self.h, first, last)

    def begin_step(self):
        lib().orc_sim_begin_step(self.h)

    def run_step_with_beam(self, dt, n_subcycles=10, do_z_push=True):
        """One time step with the beam pushed after every slice (after begin_step)."""
        lib().orc_sim_run_step_with_beam(self.h, dt, n_subcycles, int(do_z_push))

    def sort_beam(self):
        lib().orc_sim_sort_beam(self.h)

    def run_steps(self, nsteps, dt, n_subcycles=10, do_z_push=True):
        for step in range(nsteps):
            if step > 0:
                self.sort_beam()
            self.begin_step()
            self.run_step_with_beam(dt, n_subcycles, do_z_push)

    def beam(self) -> dict:
        n = self._beam[0].shape[0]
        out = {k: np.zeros(n) for k in ("x", "y", "z", "ux", "uy", "uz", "w")}
        out["valid"] = np.zeros(n, np.int32)
        out["off"] = np.zeros(self.cfg.nz + 1, np.int64)
        lib().orc_sim_get_beam(self.h, *[_dp(out[k]) for k in ("x", "y", "z", "ux", "uy", "uz", "w")],
                               out["valid"].ctypes.data_as(C.POINTER(C.c_int32)),
                               out["off"].ctypes.data_as(C.POINTER(C.c_int64)))
        return out

    def solve_slice(self, islice: int):
        lib().orc_sim_solve_slice(self.h, islice)

    def checksums(self) -> dict:
        out = np.zeros(self.ncomp)
        lib().orc_sim_checksums(self.h, _dp(out))
        return {name: out[idx] for name, idx in self.layout.items() if name != "ncomp" and idx >= 0}

    def stats(self) -> dict:
        out = np.zeros(6)
        lib().orc_sim_stats(self.h, _dp(out))
        return dict(n_qsa=out[0], mg_iters=out[1], mg_status=out[2], pc_iters=out[3], pushes=out[4],
                    np=out[5])

    def store(self) -> np.ndarray:
        out = np.zeros((self.ncomp, self.ldy, self.ldx))
        lib().orc_sim_get_store(self.h, _dp(out))
        return out

    def set_store(self, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.size == self.ncomp * self.ldx * self.ldy
        lib().orc_sim_set_store(self.h, _dp(a))

    def plasma(self) -> dict:
        n = lib().orc_sim_plasma_size(self.h)
        reals = np.zeros((11, n))
        valid = np.zeros(n, np.int32)
        lib().orc_sim_get_plasma(self.h, _dp(reals), valid.ctypes.data_as(C.POINTER(C.c_int32)))
        d = {k: reals[i] for i, k in enumerate(PLASMA_REALS)}
        d["valid"] = valid
        return d

    def set_plasma(self, d: dict):
        n = len(d["x"])
        reals = np.ascontiguousarray(np.stack([np.asarray(d[k], np.float64) for k in PLASMA_REALS]))
        valid = np.ascontiguousarray(d["valid"], dtype=np.int32)
        lib().orc_sim_set_plasma(self.h, n, _dp(reals), valid.ctypes.data_as(C.POINTER(C.c_int32)))

    # kernel-level calls on the current store / plasma
    def deposit_plasma(self, comps, ions=False) -> int:
        arr = (C.c_int * 6)(*[self.layout.get(c, -1) if isinstance(c, str) else c for c in comps])
        nq = C.c_int64(0)
        lib().orc_deposit_plasma(self.h, arr, int(ions), C.byref(nq))
        return nq.value

    def explicit_deposit(self):
        lib().orc_explicit_deposit(self.h)

    def deposit_beam(self, islice, cjx, cjy, cjz):
        f = lambda c: self.layout.get(c, -1) if isinstance(c, str) else c
        lib().orc_deposit_beam(self.h, islice, f(cjx), f(cjy), f(cjz))

    def push_plasma(self, temp_slice=False):
        lib().orc_push_plasma(self.h, int(temp_slice))

    def push_beam_slice(self, b, slice_, first, count, slip_in, dt, n_subcycles, do_z_push=True):
        """AdvanceBeamParticlesSlice on the dict of numpy arrays b (x, y, z, ux, uy, uz, w,
        valid, nsub, modified in place) with this simulation's store; returns the indices
        listed as slipped into the next slice."""
        i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        slip_in = np.ascontiguousarray(slip_in, dtype=np.int32)
        cap = b["x"].shape[0]
        out = np.zeros(max(cap, 1), np.int32)
        ext = b.get("external")
        ext = None if ext is None else np.ascontiguousarray(ext, dtype=np.float64).reshape(24)
        n = lib().orc_push_beam_slice(self.h, *[_dp(b[k]) for k in ("x", "y", "z", "ux", "uy", "uz", "w")],
                                      i32(b["valid"]), i32(b["nsub"]), b.get("charge", -1.0), b.get("mass", 1.0),
                                      first, count, i32(slip_in), slip_in.shape[0], i32(out), cap, slice_, dt,
                                      n_subcycles, int(do_z_push), None if ext is None else _dp(ext))
        return out[:n].copy()

    def solve_poisson_psi_ez_bz(self):
        lib().orc_solve_poisson_psi_ez_bz(self.h)

    def solve_poisson_bxby(self):
        lib().orc_solve_poisson_bxby(self.h)

    def initialize_sxsy_with_beam(self):
        lib().orc_initialize_sxsy_with_beam(self.h)

    def mg_solve1(self, tol_rel=1e-4, tol_abs=2.2250738585072014e-308, maxiter=200):
        it = C.c_int(0)
        st = lib().orc_mg_solve1_store(self.h, tol_rel, tol_abs, maxiter, C.byref(it))
        return st, it.value


def dst1(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib().orc_dst1(_dp(x), _dp(out), x.shape[0])
    return out


def poisson(cfg, staging: np.ndarray) -> np.ndarray:
    c = make_config(cfg)
    st = np.ascontiguousarray(staging, dtype=np.float64).reshape(cfg.ny, cfg.nx)
    lhs = np.zeros((cfg.ny + 4, cfg.nx + 4))
    lib().orc_poisson(C.byref(c), _dp(st), _dp(lhs))
    return lhs


def bf16_round(x: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 with round-to-nearest-even, returned as FP32 values (finite inputs)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def split_bf16(v: np.ndarray):
    """v (rounded to FP32) as hi + lo, hi = bf16(v), lo = bf16(v - hi)."""
    v = np.asarray(v, dtype=np.float32)
    hi = bf16_round(v)
    return hi, bf16_round(v - hi)


def _gemm_t_bf16x3(a: np.ndarray, s_hi: np.ndarray, s_lo: np.ndarray) -> np.ndarray:
    """(a S)^T with a split into BF16 hi/lo: hi*hi + hi*lo + lo*hi, sums exact, rounded to FP32."""
    ah, al = split_bf16(a)
    ah, al = ah.astype(np.float64), al.astype(np.float64)
    d = ah @ s_hi + ah @ s_lo + al @ s_hi
    return d.astype(np.float32).T


def poisson_bf16x3(cfg, staging: np.ndarray) -> np.ndarray:
    """The split-BF16 sine-matrix Dirichlet solve (poisson_solver="bf16x3") restated in numpy:
    F = Sy (L o (Sy X Sx)) Sx with S_jk = 2 sin(pi (j+1)(k+1)/(n+1)) (RODFT00,
    FFTPoissonSolverDirichletDirect.cpp:87-139) and L the eigenvalue matrix with the
    normalisation 1/(4 (nx+1)(ny+1)) (:60-82), evaluated as four products
    out = (in S)^T whose operands are split into two BF16 terms each. Products are summed
    exactly here and rounded to FP32 per product; returns the valid ny x nx solution."""
    nx, ny = cfg.nx, cfg.ny
    X = np.asarray(staging, dtype=np.float64).reshape(ny, nx)
    def sine(n):
        k = np.arange(1, n + 1)
        return 2.0 * np.sin(np.pi * np.outer(k, k) / (n + 1))
    sx, sy = sine(nx), sine(ny)
    # S is split from its FP64 value: hi = bf16(S), lo = bf16(S - hi)
    sxh = bf16_round(sx).astype(np.float64)
    sxl = bf16_round(sx - sxh).astype(np.float64)
    syh = bf16_round(sy).astype(np.float64)
    syl = bf16_round(sy - syh).astype(np.float64)
    dx, dy = cfg.dx, cfg.dy
    fx = np.sin(np.arange(1, nx + 1) * np.pi / (2.0 * (nx + 1))) ** 2
    fy = np.sin(np.arange(1, ny + 1) * np.pi / (2.0 * (ny + 1))) ** 2
    ev = (0.5 / (2 * ((nx + 1) * (ny + 1)))) / (-4.0 * (fx[:, None] / dx ** 2 + fy[None, :] / dy ** 2))
    p1 = _gemm_t_bf16x3(X, sxh, sxl)
    p2 = (_gemm_t_bf16x3(p1, syh, syl).astype(np.float64) * ev.T).astype(np.float32)
    p3 = _gemm_t_bf16x3(p2, sxh, sxl)
    return _gemm_t_bf16x3(p3, syh, syl).astype(np.float64)


def shape(order, xmid, ix):
    out = np.zeros(2)
    lib().orc_shape(order, xmid, ix, _dp(out))
    return out[0], int(out[1])


def dshape(dtype, order, xmid, ix):
    out = np.zeros(3)
    lib().orc_dshape(dtype, order, xmid, ix, _dp(out))
    return out[0], out[1], int(out[2])


# ---- tile binning (integer work, numpy) ----------------------------------------------

def start_cell(order: int, xmid: np.ndarray) -> np.ndarray:
    """Lowest cell of the order-`order` deposition stencil: the cell of index 0 of
    compute_single_shape_factor (src/particles/deposition/ShapeFactors.H:122-195), i.e.
    floor(xmid + 0.5) - 1 for order 2, floor(xmid) - 1 for order 3, floor(xmid) for order 1,
    floor(xmid + 0.5) for order 0."""
    xmid = np.asarray(xmid, dtype=np.float64)
    if order == 0:
        return np.floor(xmid + 0.5).astype(np.int64)
    if order == 1:
        return np.floor(xmid).astype(np.int64)
    if order == 2:
        return np.floor(xmid + 0.5).astype(np.int64) - 1
    return np.floor(xmid).astype(np.int64) - 1


def tile_keys(cfg, x, y, valid, tile: int):
    """Bin key of every particle for the CPU tiled deposition: get_start_cell
    (PlasmaDepositCurrent.cpp:137-155) of the particle, then
    ((cx - lo) / T) * ntile_y + (cy - lo) / T with lo = -2 (the first guard cell) and C
    integer division; invalid particles get ntile_x * ntile_y
    (src/particles/deposition/DepositionUtil.H:213-227). Returns (keys, nbins)."""
    dx = (cfg.xhi - cfg.xlo) / cfg.nx
    dy = (cfg.yhi - cfg.ylo) / cfg.ny
    xoff = 0.5 * (cfg.xlo + cfg.xhi - dx * (cfg.nx - 1))   # GetPosOffset, Fields.H:72-77
    yoff = 0.5 * (cfg.ylo + cfg.yhi - dy * (cfg.ny - 1))
    xmid = (np.asarray(x, dtype=np.float64) - xoff) * (1.0 / dx)
    ymid = (np.asarray(y, dtype=np.float64) - yoff) * (1.0 / dy)
    ntx = (cfg.nx + 4 + tile - 1) // tile
    nty = (cfg.ny + 4 + tile - 1) // tile
    cx = start_cell(cfg.depos_order, xmid) + 2
    cy = start_cell(cfg.depos_order, ymid) + 2
    tx = np.clip(np.trunc(cx / tile).astype(np.int64), 0, ntx - 1)
    ty = np.clip(np.trunc(cy / tile).astype(np.int64), 0, nty - 1)
    keys = tx * nty + ty
    keys[np.asarray(valid) == 0] = ntx * nty
    return keys, ntx * nty + 1


def dense_bins(keys: np.ndarray, nbins: int):
    """amrex::DenseBins::build as a counting sort: histogram of the keys, exclusive scan,
    then every particle, in index order, goes to the next free slot of its bin.
    Returns (perm, offsets[nbins + 1])."""
    keys = np.asarray(keys, dtype=np.int64)
    counts = np.bincount(keys, minlength=nbins)
    offsets = np.zeros(nbins + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(counts)
    nxt = offsets[:-1].copy()
    perm = np.empty(keys.size, dtype=np.int64)
    for ip, k in enumerate(keys.tolist()):
        perm[nxt[k]] = ip
        nxt[k] += 1
    return perm, offsets
