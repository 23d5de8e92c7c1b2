"""Run configuration and input generation for the per-slice quasi-static loop.

`SliceConfig` carries the reference's input-deck keys that this path honours
(amr.n_cell, geometry.prob_lo/hi, hipace.depos_order_xy, hipace.depos_derivative_type,
hipace.bxby_solver, hipace.MG_tolerance_rel/abs, hipace.predcorr_*, boundary.particle,
plasmas.*, max_qsa_weighting_factor; src/Hipace.cpp:97-213, src/Hipace.H:208-252,
src/particles/plasma/PlasmaParticleContainer.cpp:30-170).

`parse_deck` reads the reference's own input-deck text format (``key = value``
lines, ``#`` comments) for the keys above and the beam keys of
src/particles/beam/BeamParticleContainer.cpp:42-200, so the reference's example
decks and the overrides of its test scripts can be run unchanged.

Beams are generated on the host (numpy) and handed to the slice loop sorted into
ζ-slices (compressed row offsets), the way BoxSorter::sortParticlesByBox
(src/particles/sorting/BoxSort.cpp:14-78) hands them to the per-slice tiles.
"""
from __future__ import annotations

import ast
import dataclasses
import math
import re
import shlex
from typing import Dict, Optional, Tuple

import numpy as np

DBL_MIN = 2.2250738585072014e-308

BC_CODES = {"Reflecting": 0, "Periodic": 1, "Absorbing": 2}


@dataclasses.dataclass
class SliceConfig:
    nx: int
    ny: int
    nz: int
    xlo: float
    xhi: float
    ylo: float
    yhi: float
    zlo: float
    zhi: float
    depos_order: int = 2
    derivative_type: int = 2
    explicit_solver: bool = True
    deposit_rho: bool = False
    do_beam_jx_jy: bool = True
    mg_tol_rel: float = 1e-4
    mg_tol_abs: float = DBL_MIN
    mg_maxiter: int = 200
    particle_bc: str = "Periodic"
    particle_lo: Optional[Tuple[float, float]] = None
    particle_hi: Optional[Tuple[float, float]] = None
    pc_tol: float = 4e-2
    pc_maxiter: int = 30
    pc_mix: float = 0.05
    has_plasma: bool = True
    ppc: Tuple[int, int] = (1, 1)
    density: float = 1.0
    ramp_a: float = 0.0
    # <plasma>.density(x,y,z) when it is an expression rather than a number: evaluated at every
    # lattice position (z = 0, the start of the run) instead of density * (1 + ramp_a r^2)
    density_expr: Optional[str] = None
    plasma_radius: float = math.inf
    u_mean: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    plasma_charge: float = -1.0
    plasma_mass: float = 1.0
    max_qsa: float = 35.0
    neutralize: bool = True
    # order of the FP64 deposition sums: "ordered" (the reference CPU build's tiled order, bit
    # for bit reproducible) or "atomic" (LDS windows + global atomics, hardware order)
    deposition: str = "ordered"
    # hipace.tile_size (src/Hipace.cpp:160): tiles of the CPU build's deposition order
    tile_size: int = 32
    # atomic deposition: walks the particles sorted by 32x8-cell tiles, re-sorted every
    # bin_period slices (0: index order); only the order of floating-point sums depends on it
    bin_period: int = 4
    # plasmas.reorder_period (PlasmaParticleContainer.cpp:197-208): reorder the plasma arrays
    # in place by deposition tile every reorder_period slices (atomic deposition only; 0: never)
    reorder_period: int = 0
    # "fp64": the drop-in path; "mixed": FP32 gather, push increments and deposition with
    # FP32 contributions summed in FP64 LDS words (FP64 store, particle state and multigrid)
    precision: str = "fp64"
    # the mixed path's window sums: "fp64" (FP32 contributions into FP64 LDS words) or
    # "kahan32" (compensated FP32 sums: an FP32 sum word and an FP32 error word per cell)
    mixed_sums: str = "fp64"
    # the explicit solver's multigrid starts from "previous" (the previous slice's Bx, By, as
    # the reference does), "linear" (extrapolated from the two previous slices) or "dst" (the
    # previous slice's Bx, By corrected mg_guess_iters times by the shifted DST solve with
    # chi replaced by its mean)
    mg_guess: str = "previous"
    mg_guess_iters: int = 1
    # transverse Poisson solves: "fft" (DST-I by FFT, FP64) or "bf16x3" (DST-I as sine-matrix
    # GEMMs on the matrix cores, FP32 operands split into two BF16 terms each)
    poisson_solver: str = "fft"
    # time step and beam push (hipace.dt; <beam>.n_subcycles, <beam>.do_z_push,
    # src/particles/beam/BeamParticleContainer.cpp:42-90)
    dt: float = 0.0
    n_subcycles: int = 10
    do_z_push: bool = True
    # SALAME (src/Hipace.cpp:137-143): a beam has do_salame; hipace.salame_n_iter,
    # salame_do_advance, salame_relative_tolerance, salame_Ez_target(zeta,zeta_initial,Ez_initial)
    salame: bool = False
    salame_n_iter: int = 5
    salame_do_advance: bool = True
    salame_relative_tolerance: float = 1e-4
    salame_target: str = "Ez_initial"

    @property
    def dx(self) -> float:
        return (self.xhi - self.xlo) / self.nx

    @property
    def dy(self) -> float:
        return (self.yhi - self.ylo) / self.ny

    @property
    def dz(self) -> float:
        return (self.zhi - self.zlo) / self.nz

    @property
    def bounds(self):
        lo = self.particle_lo or (self.xlo, self.ylo)
        hi = self.particle_hi or (self.xhi, self.yhi)
        return lo, hi

    def slice_z(self, islice: int) -> float:
        """ζ of the centre of slice `islice` (GetPosOffset, src/fields/Fields.H:72-77)."""
        return self.zlo + (islice + 0.5) * self.dz


@dataclasses.dataclass
class BeamSpec:
    injection_type: str = "fixed_ppc"      # fixed_ppc | fixed_weight
    profile: str = "flattop"               # flattop | gaussian (fixed_ppc); gaussian (fixed_weight)
    density: float = 0.0
    zmin: float = -math.inf
    zmax: float = math.inf
    radius: float = math.inf
    position_mean: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    position_std: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    ppc: Tuple[int, int, int] = (1, 1, 1)
    u_mean: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    u_std: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    num_particles: int = 0
    charge: float = -1.0
    mass: float = 1.0
    seed: int = 0
    min_density: float = 0.0
    # beams.external_E/B(x,y,z,t) as (6, 4) affine coefficients (Ex, Ey, Ez, Bx, By, Bz) x
    # (1, x, y, z), or None
    external: Optional[np.ndarray] = None
    do_salame: bool = False


@dataclasses.dataclass
class BeamArrays:
    """Beam particles grouped by ζ-slice: particles of slice s are [off[s], off[s+1])."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    ux: np.ndarray
    uy: np.ndarray
    uz: np.ndarray
    w: np.ndarray
    valid: np.ndarray
    off: np.ndarray
    charge: float = -1.0
    mass: float = 1.0
    external: Optional[np.ndarray] = None   # (6, 4) affine external fields, or None
    salame: Optional[np.ndarray] = None     # int32 per particle: 1 for a SALAME beam's particles

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @staticmethod
    def empty(nz: int) -> "BeamArrays":
        z = np.zeros(0)
        return BeamArrays(z, z.copy(), z.copy(), z.copy(), z.copy(), z.copy(), z.copy(),
                          np.zeros(0, np.int32), np.zeros(nz + 1, np.int64))


# --------------------------------------------------------------------------------------
# beam generation
# --------------------------------------------------------------------------------------

def _enforce_bc_scalar(cfg: SliceConfig, x, y, ux, uy):
    """EnforceBC for beam particles at injection (src/particles/pusher/GetAndSetPosition.H:29-99)."""
    (plx, ply), (phx, phy) = cfg.bounds
    out = (x < plx) | (y < ply) | (x > phx) | (y > phy)
    valid = np.ones_like(x, dtype=bool)
    if not out.any():
        return x, y, ux, uy, valid
    lx, ly = phx - plx, phy - ply
    if cfg.particle_bc == "Periodic":
        x = np.where(out, np.fmod(x - plx, lx), x)
        x = np.where(out & (x < 0), x + lx, x)
        x = np.where(out, x + plx, x)
        y = np.where(out, np.fmod(y - ply, ly), y)
        y = np.where(out & (y < 0), y + ly, y)
        y = np.where(out, y + ply, y)
    elif cfg.particle_bc == "Reflecting":
        xr = np.fmod(x - plx, 2 * lx)
        xr = np.where(xr < 0, xr + 2 * lx, xr) + plx
        fx = xr > phx
        xr = np.where(fx, 2 * phx - xr, xr)
        ux = np.where(out & fx, -ux, ux)
        x = np.where(out, xr, x)
        yr = np.fmod(y - ply, 2 * ly)
        yr = np.where(yr < 0, yr + 2 * ly, yr) + ply
        fy = yr > phy
        yr = np.where(fy, 2 * phy - yr, yr)
        uy = np.where(out & fy, -uy, uy)
        y = np.where(out, yr, y)
    else:
        valid = ~out
    return x, y, ux, uy, valid


def make_fixed_ppc_beam(cfg: SliceConfig, spec: BeamSpec) -> BeamArrays:
    """Deterministic beam, one slice at a time, following
    BeamParticleContainer::InitBeamFixedPPCSlice (src/particles/beam/
    BeamParticleContainerInit.cpp:198-346) with even particle spacing
    (get_position_unit_cell, src/particles/particles_utils/ParticleUtil.H:45-61)
    and the flattop/gaussian density of GetInitialDensity.H:31-49."""
    nx, ny = cfg.nx, cfg.ny
    dx, dy, dz = cfg.dx, cfg.dy, cfg.dz
    px, py, pz = spec.ppc
    nppc = px * py * pz
    scale_fac = 1.0 / nppc
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")  # x fastest
    ii = ii.ravel().astype(np.float64)
    jj = jj.ravel().astype(np.float64)
    parts = {k: [] for k in ("x", "y", "z", "w")}
    off = np.zeros(cfg.nz + 1, np.int64)
    xm, ym, zm = spec.position_mean
    xs, ys, zs = spec.position_std
    for isl in range(cfg.nz):
        xs_l, ys_l, zs_l, ws_l = [], [], [], []
        for ipart in range(nppc):
            ixp = ipart // (py * pz)
            iyp = (ipart % (py * pz)) % py
            izp = (ipart % (py * pz)) // py
            r0 = (0.5 + ixp) / px
            r1 = (0.5 + iyp) / py
            r2 = (0.5 + izp) / pz
            x = cfg.xlo + (ii + r0) * dx
            y = cfg.ylo + (jj + r1) * dy
            z = cfg.zlo + (isl + r2) * dz
            keep = ~((z >= spec.zmax) | (z < spec.zmin) |
                     (((x - xm) * (x - xm) + (y - ym) * (y - ym)) > spec.radius * spec.radius))
            if spec.profile == "gaussian":
                dens = (spec.density * np.exp(-0.5 * ((x - xm) / xs) ** 2)
                        * np.exp(-0.5 * ((y - ym) / ys) ** 2)
                        * math.exp(-0.5 * ((z - zm) / zs) ** 2))
            else:
                dens = np.full_like(x, abs(spec.density))
            keep &= ~(dens <= spec.min_density)
            xs_l.append(np.where(keep, x, np.nan))
            ys_l.append(np.where(keep, y, np.nan))
            zs_l.append(np.full_like(x, z))
            ws_l.append(np.where(keep, dens * scale_fac, np.nan))
        # order: cells (x fastest), then i_part within a cell
        X = np.stack(xs_l, axis=1).ravel()
        Y = np.stack(ys_l, axis=1).ravel()
        Z = np.stack(zs_l, axis=1).ravel()
        W = np.stack(ws_l, axis=1).ravel()
        m = ~np.isnan(X)
        parts["x"].append(X[m])
        parts["y"].append(Y[m])
        parts["z"].append(Z[m])
        parts["w"].append(W[m])
        off[isl + 1] = off[isl] + int(m.sum())
    x = np.concatenate(parts["x"]) if parts["x"] else np.zeros(0)
    y = np.concatenate(parts["y"]) if parts["y"] else np.zeros(0)
    z = np.concatenate(parts["z"]) if parts["z"] else np.zeros(0)
    w = np.abs(np.concatenate(parts["w"])) if parts["w"] else np.zeros(0)
    n = x.shape[0]
    ux = np.full(n, spec.u_mean[0] * 1.0)
    uy = np.full(n, spec.u_mean[1] * 1.0)
    uz = np.full(n, spec.u_mean[2] * 1.0)
    x, y, ux, uy, valid = _enforce_bc_scalar(cfg, x, y, ux, uy)
    return BeamArrays(x, y, z, ux, uy, uz, w, valid.astype(np.int32), off, spec.charge, spec.mass)


def make_fixed_weight_beam(cfg: SliceConfig, spec: BeamSpec) -> BeamArrays:
    """Gaussian beam of `num_particles` equal-weight macro-particles drawn from a seeded
    Philox generator (the reference's AMReX RNG stream cannot be reproduced; the same
    arrays are fed to every path). Weight per BeamParticleContainer::InitData
    (src/particles/beam/BeamParticleContainer.cpp:179-189) and InitBeamFixedWeightSlice
    (BeamParticleContainerInit.cpp:382-475): w = n_peak (2π)^{3/2} σxσyσz / (dx dy dz N)."""
    rng = np.random.Generator(np.random.Philox(spec.seed))
    n = int(spec.num_particles)
    z = rng.normal(spec.position_mean[2], spec.position_std[2], n)
    x = rng.normal(0.0, spec.position_std[0], n)
    y = rng.normal(0.0, spec.position_std[1], n)
    ux = spec.u_mean[0] + rng.normal(0.0, 1.0, n) * spec.u_std[0]
    uy = spec.u_mean[1] + rng.normal(0.0, 1.0, n) * spec.u_std[1]
    uz = spec.u_mean[2] + rng.normal(0.0, 1.0, n) * spec.u_std[2]
    valid = ~((z < spec.zmin) | (z > spec.zmax) | (x * x + y * y > spec.radius * spec.radius))
    x = x + spec.position_mean[0]
    y = y + spec.position_mean[1]
    total = spec.density * spec.charge
    for s in spec.position_std:
        total *= s * math.sqrt(2.0 * math.pi)
    total *= (1.0 / cfg.dx) * (1.0 / cfg.dy) * (1.0 / cfg.dz)
    weight = abs(total / (n * spec.charge)) if n else 0.0
    w = np.full(n, weight)
    x, y, ux, uy, v2 = _enforce_bc_scalar(cfg, x, y, ux, uy)
    valid &= v2
    # sort into ζ-boxes: box = int((z - zlo) / dz), outside -> dropped (BoxSort.cpp:40-45)
    box = np.floor((z - cfg.zlo) * (1.0 / cfg.dz)).astype(np.int64)
    inside = (box >= 0) & (box < cfg.nz)
    order = np.argsort(np.where(inside, box, cfg.nz), kind="stable")
    order = order[inside[order]]
    counts = np.bincount(box[inside], minlength=cfg.nz)
    off = np.zeros(cfg.nz + 1, np.int64)
    off[1:] = np.cumsum(counts)
    return BeamArrays(x[order], y[order], z[order], ux[order], uy[order], uz[order], w[order],
                      valid[order].astype(np.int32), off, spec.charge, spec.mass)


def make_beam(cfg: SliceConfig, spec: Optional[BeamSpec]) -> BeamArrays:
    if spec is None:
        return BeamArrays.empty(cfg.nz)
    if spec.injection_type == "fixed_ppc":
        b = make_fixed_ppc_beam(cfg, spec)
    elif spec.injection_type == "fixed_weight":
        b = make_fixed_weight_beam(cfg, spec)
    else:
        raise ValueError(f"unsupported beam injection_type {spec.injection_type!r}")
    b.external = None if spec.external is None else np.array(spec.external, dtype=np.float64)
    if spec.do_salame:
        b.salame = np.ones(b.n, np.int32)
    return b


def merge_beams(cfg: SliceConfig, beams) -> BeamArrays:
    """Concatenate several species slice by slice (all must share charge and mass)."""
    beams = [b for b in beams if b is not None]
    if not beams:
        return BeamArrays.empty(cfg.nz)
    if len(beams) == 1:
        return beams[0]
    ext = [b.external for b in beams]
    if any(e is not None for e in ext) and not all(e is not None and np.array_equal(e, ext[0]) for e in ext):
        raise ValueError("merged beams must share their external fields")
    fields = ("x", "y", "z", "ux", "uy", "uz", "w", "valid")
    out = {f: [] for f in fields}
    sal = []
    off = np.zeros(cfg.nz + 1, np.int64)
    for s in range(cfg.nz):
        cnt = 0
        for b in beams:
            a, e = int(b.off[s]), int(b.off[s + 1])
            for f in fields:
                out[f].append(getattr(b, f)[a:e])
            sal.append(b.salame[a:e] if b.salame is not None else np.zeros(e - a, np.int32))
            cnt += e - a
        off[s + 1] = off[s] + cnt
    merged = BeamArrays(*(np.concatenate(out[f]) for f in fields), off, beams[0].charge, beams[0].mass, ext[0])
    if any(b.salame is not None for b in beams):
        merged.salame = np.concatenate(sal).astype(np.int32)
    return merged


# --------------------------------------------------------------------------------------
# input decks
# --------------------------------------------------------------------------------------

def _tokens(val: str):
    return shlex.split(val, comments=False)


def parse_kv(text: str) -> Dict[str, str]:
    out: Dict[str, str] = {}
    for line in text.splitlines():
        line = line.split("#", 1)[0].strip()
        if not line or "=" not in line:
            continue
        k, v = line.split("=", 1)
        out[k.strip()] = v.strip()
    return out


def _num(v: str) -> float:
    v = v.strip().strip('"')
    if re.fullmatch(r"[-+0-9.eE]+", v):
        return float(v)
    raise ValueError(f"only numeric values are supported here, got {v!r}")


_BINOPS = {ast.Add: lambda a, b: a + b, ast.Sub: lambda a, b: a - b}


def _affine(node, names):
    """An expression of the AMReX parser's arithmetic as an affine form: array
    (c, d_x, d_y, d_z, d_t) with value c + d_x x + d_y y + d_z z + d_t t. Products and
    quotients need a constant factor (or divisor); powers need constant operands."""
    if isinstance(node, ast.Expression):
        return _affine(node.body, names)
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return np.array([float(node.value), 0, 0, 0, 0])
    if isinstance(node, ast.Name) and node.id in names:
        v = np.zeros(5)
        v[1 + names.index(node.id)] = 1.0
        return v
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _affine(node.operand, names)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.BinOp):
        a, b = _affine(node.left, names), _affine(node.right, names)
        const = lambda v: not np.any(v[1:])
        if type(node.op) in _BINOPS:
            return _BINOPS[type(node.op)](a, b)
        if isinstance(node.op, ast.Mult):
            if const(a):
                return a[0] * b
            if const(b):
                return a * b[0]
        if isinstance(node.op, ast.Div) and const(b):
            return a / b[0]
        if isinstance(node.op, ast.Pow) and const(a) and const(b):
            return np.array([a[0] ** b[0], 0, 0, 0, 0])
        raise ValueError("external fields must be affine in x, y, z on this path")
    raise ValueError(f"unsupported term in an external-field expression: {ast.dump(node)}")


def parse_affine(expr: str) -> np.ndarray:
    """(c0, cx, cy, cz) of an affine, time-independent function of (x, y, z) written in the
    reference's parser syntax (`^` for powers)."""
    tree = ast.parse(expr.strip().strip('"').replace("^", "**"), mode="eval")
    v = _affine(tree, ["x", "y", "z", "t"])
    if v[4] != 0.0:
        raise ValueError("time-dependent external fields are not on this path")
    return v[:4]


def external_fields(e_exprs, b_exprs) -> Optional[np.ndarray]:
    """(6, 4) coefficients of beams.external_E(x,y,z,t) and external_B(x,y,z,t)
    (src/particles/beam/BeamParticleContainer.cpp:72-87); None when neither is given."""
    if e_exprs is None and b_exprs is None:
        return None
    out = np.zeros((6, 4))
    for k, exprs in ((0, e_exprs), (3, b_exprs)):
        if exprs is None:
            continue
        if len(exprs) != 3:
            raise ValueError("external fields take three components")
        for q in range(3):
            out[k + q] = parse_affine(exprs[q])
    return out


_FUNCS = {"sqrt": math.sqrt, "exp": math.exp, "log": math.log, "sin": math.sin, "cos": math.cos,
          "tan": math.tan, "abs": abs, "min": min, "max": max, "pow": pow}


def parse_function(expr: str, names):
    """A function of `names` from an expression in the reference's parser syntax (numbers,
    the names, + - * / and ^, parentheses and a few elementary functions), evaluated by
    walking the syntax tree; nothing else is accepted."""
    tree = ast.parse(expr.strip().strip('"').replace("^", "**"), mode="eval")
    ops = {ast.Add: lambda a, b: a + b, ast.Sub: lambda a, b: a - b, ast.Mult: lambda a, b: a * b,
           ast.Div: lambda a, b: a / b, ast.Pow: lambda a, b: a ** b}

    def ev(node, env):
        if isinstance(node, ast.Expression):
            return ev(node.body, env)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return float(node.value)
        if isinstance(node, ast.Name) and node.id in env:
            return env[node.id]
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = ev(node.operand, env)
            return -v if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.BinOp) and type(node.op) in ops:
            return ops[type(node.op)](ev(node.left, env), ev(node.right, env))
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in _FUNCS \
                and not node.keywords:
            return float(_FUNCS[node.func.id](*[ev(a, env) for a in node.args]))
        raise ValueError(f"unsupported term in {expr!r}")

    # every name must be one of the arguments: check once on the tree
    for node in ast.walk(tree):
        if isinstance(node, ast.Name) and node.id not in names and node.id not in _FUNCS:
            raise ValueError(f"unknown name {node.id!r} in {expr!r}")
    return lambda *args: ev(tree, dict(zip(names, (float(a) for a in args))))


_NP_FUNCS = {"sqrt": np.sqrt, "exp": np.exp, "log": np.log, "sin": np.sin, "cos": np.cos, "tan": np.tan,
             "abs": np.abs, "min": np.minimum, "max": np.maximum, "pow": np.power}


def parse_function_np(expr: str, names):
    """parse_function for numpy arrays: the same syntax and checks, elementwise evaluation in
    the tree's order (every operation rounds as the scalar evaluation does)."""
    parse_function(expr, names)
    tree = ast.parse(expr.strip().strip('"').replace("^", "**"), mode="eval")
    ops = {ast.Add: np.add, ast.Sub: np.subtract, ast.Mult: np.multiply, ast.Div: np.divide, ast.Pow: np.power}

    def ev(node, env):
        if isinstance(node, ast.Expression):
            return ev(node.body, env)
        if isinstance(node, ast.Constant):
            return np.float64(node.value)
        if isinstance(node, ast.Name):
            return env[node.id]
        if isinstance(node, ast.UnaryOp):
            v = ev(node.operand, env)
            return np.negative(v) if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.BinOp):
            return ops[type(node.op)](ev(node.left, env), ev(node.right, env))
        return _NP_FUNCS[node.func.id](*[ev(a, env) for a in node.args])

    return lambda *args: ev(tree, dict(zip(names, (np.asarray(a, dtype=np.float64) for a in args))))


def lattice_density(cfg: "SliceConfig", z: float = 0.0) -> np.ndarray:
    """<plasma>.density(x,y,z) at every lattice position of the injection, in the layout of
    hpq_plasma_profile.density_table: entry (j*nx + i)*nppc + i_part, i_part = ix + iy*ppc_x,
    x = xlo + (i + (0.5 + ix)/ppc_x)*dx (ParticleUtil.H:65-104; the same rounding as the
    injection kernel)."""
    px, py = cfg.ppc
    f = parse_function_np(cfg.density_expr, ["x", "y", "z"])
    ip = np.arange(px * py)
    r0 = (0.5 + (ip % px)) / px
    r1 = (0.5 + (ip // px)) / py
    i = np.arange(cfg.nx, dtype=np.float64)
    j = np.arange(cfg.ny, dtype=np.float64)
    x = cfg.xlo + (i[None, :, None] + r0[None, None, :] + 0.0) * cfg.dx
    y = cfg.ylo + (j[:, None, None] + r1[None, None, :] + 0.0) * cfg.dy
    x, y = np.broadcast_arrays(x, y)
    return np.ascontiguousarray(np.broadcast_to(f(x, y, np.full(x.shape, z)), x.shape)).reshape(-1)


def parse_deck(text: str, overrides: Optional[Dict[str, str]] = None):
    """Parse a reference input deck into (SliceConfig, [BeamSpec], max_step)."""
    kv = parse_kv(text)
    if overrides:
        kv.update({k.strip(): str(v).strip() for k, v in overrides.items()})
    n = [int(_num(t)) for t in _tokens(kv["amr.n_cell"])]
    lo = [_num(t) for t in _tokens(kv["geometry.prob_lo"])]
    hi = [_num(t) for t in _tokens(kv["geometry.prob_hi"])]
    if kv.get("boundary.field", "Dirichlet") != "Dirichlet":
        raise ValueError("only boundary.field = Dirichlet is on this path")
    cfg = SliceConfig(nx=n[0], ny=n[1], nz=n[2], xlo=lo[0], xhi=hi[0], ylo=lo[1], yhi=hi[1],
                      zlo=lo[2], zhi=hi[2])
    cfg.depos_order = int(_num(kv.get("hipace.depos_order_xy", "2")))
    cfg.derivative_type = int(_num(kv.get("hipace.depos_derivative_type", "2")))
    cfg.explicit_solver = kv.get("hipace.bxby_solver", "explicit").strip('"') == "explicit"
    cfg.mg_tol_rel = _num(kv.get("hipace.MG_tolerance_rel", "1e-4"))
    if "hipace.MG_tolerance_abs" in kv:
        cfg.mg_tol_abs = _num(kv["hipace.MG_tolerance_abs"])
    cfg.pc_tol = _num(kv.get("hipace.predcorr_B_error_tolerance", "4e-2"))
    cfg.pc_maxiter = int(_num(kv.get("hipace.predcorr_max_iterations", "30")))
    cfg.pc_mix = _num(kv.get("hipace.predcorr_B_mixing_factor", "0.05"))
    cfg.do_beam_jx_jy = _num(kv.get("hipace.do_beam_jx_jy_deposition", "1")) != 0
    cfg.tile_size = int(_num(kv.get("hipace.tile_size", "32")))
    cfg.particle_bc = kv.get("boundary.particle", "Absorbing").strip('"')
    fd = kv.get("diagnostic.field_data", "")
    cfg.deposit_rho = ("rho" in _tokens(fd)) or _num(kv.get("hipace.deposit_rho", "0")) != 0
    pnames = _tokens(kv.get("plasmas.names", "no_plasma"))
    if pnames and pnames != ["no_plasma"]:
        if len(pnames) != 1:
            raise ValueError("one plasma species on this path")
        p = pnames[0]
        dens = kv.get(f"{p}.density(x,y,z)", kv.get(f"{p}.density", "1.")).strip('"')
        try:
            cfg.density = _num(dens)
        except ValueError:
            parse_function(dens, ["x", "y", "z"])   # rejects anything but the parser syntax
            cfg.density_expr = dens
        cfg.ppc = tuple(int(_num(t)) for t in _tokens(kv.get(f"{p}.ppc", "1 1")))
        um = kv.get(f"{p}.u_mean", "0 0 0")
        cfg.u_mean = tuple(_num(t) for t in _tokens(um))
        el = kv.get(f"{p}.element", "electron")
        if el != "electron":
            raise ValueError("electron plasma only on this path")
        cfg.max_qsa = _num(kv.get(f"{p}.max_qsa_weighting_factor",
                                  kv.get("plasmas.max_qsa_weighting_factor", "35")))
        cfg.reorder_period = int(_num(kv.get(f"{p}.reorder_period", kv.get("plasmas.reorder_period", "0"))))
        cfg.has_plasma = True
    else:
        cfg.has_plasma = False
    beams = []
    for b in _tokens(kv.get("beams.names", "no_beam")):
        if b == "no_beam":
            continue
        g = lambda key, d=None: kv.get(f"{b}.{key}", kv.get(f"beams.{key}", d))
        spec = BeamSpec()
        spec.injection_type = g("injection_type")
        spec.profile = g("profile", "gaussian" if spec.injection_type == "fixed_weight" else "flattop")
        spec.density = abs(_num(g("density", "0")))
        spec.zmin = _num(g("zmin")) if g("zmin") else -math.inf
        spec.zmax = _num(g("zmax")) if g("zmax") else math.inf
        spec.radius = _num(g("radius")) if g("radius") else math.inf
        if g("position_mean"):
            spec.position_mean = tuple(_num(t) for t in _tokens(g("position_mean")))
        if g("position_std"):
            spec.position_std = tuple(_num(t) for t in _tokens(g("position_std")))
        if g("ppc"):
            spec.ppc = tuple(int(_num(t)) for t in _tokens(g("ppc")))
        spec.u_mean = tuple(_num(t) for t in _tokens(g("u_mean", "0 0 0")))
        spec.u_std = tuple(_num(t) for t in _tokens(g("u_std", "0 0 0")))
        if g("num_particles"):
            spec.num_particles = int(_num(g("num_particles")))
        if g("element", "electron") != "electron":
            raise ValueError("electron beams only on this path")
        cfg.n_subcycles = int(_num(g("n_subcycles", "10")))
        spec.do_salame = _num(g("do_salame", "0")) != 0
        cfg.do_z_push = _num(g("do_z_push", "1")) != 0
        ext_e, ext_b = g("external_E(x,y,z,t)"), g("external_B(x,y,z,t)")
        spec.external = external_fields(None if ext_e is None else _tokens(ext_e),
                                        None if ext_b is None else _tokens(ext_b))
        beams.append(spec)
    cfg.dt = _num(kv.get("hipace.dt", "0"))
    cfg.salame = any(b.do_salame for b in beams)
    cfg.salame_n_iter = int(_num(kv.get("hipace.salame_n_iter", "5")))
    cfg.salame_do_advance = _num(kv.get("hipace.salame_do_advance", "1")) != 0
    cfg.salame_relative_tolerance = _num(kv.get("hipace.salame_relative_tolerance", "1e-4"))
    cfg.salame_target = kv.get("hipace.salame_Ez_target(zeta,zeta_initial,Ez_initial)", "Ez_initial").strip('"')
    max_step = int(_num(kv.get("max_step", "0")))
    return cfg, beams, max_step
