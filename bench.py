"""Throughput of the per-slice quasi-static loop on one MI355X.

Workload (one step = one pass of the slice loop over all ζ-slices, inputs resident in HBM):
  511 x 511 transverse grid over [-8, 8]^2, 512 slices over ζ in [-8, 4], uniform plasma
  n0 = 1 at 2 x 2 particles per cell (1 044 484 plasma particles per slice), shape order 2,
  explicit Bx/By solver with multigrid; Gaussian electron drive beam with peak density
  0.01 n0, sigma_r = 0.3, sigma_z = 1.41, 4e6 macro-particles (seeded Philox), gamma = 1000.

Three paths are timed on the same workload: the FP64 plain path (the drop-in: deposition
summed in the reference CPU build's order, bit for bit reproducible; the headline `value`),
the FP64 path with atomic deposition sums, and the mixed path (FP32 gather, push increments and
deposition, the FP32 contributions summed in FP64 LDS words; DST-I as split-BF16 sine-matrix
GEMMs on the matrix cores; FP64 multigrid and particle state). One more untimed step of each
records Ez, Ex-By, jx and rho-jz on every slice, and the other paths' norm-wise relative
errors against the FP64 path are reported.

Prints one JSON line: slices/s of the FP64 path (value), plasma particle pushes/s, the roofline
of the kernel family that takes the most time and the achieved bandwidth of every family (HIP-
event timed), and the oracle timed on the host cores (all cores of the box's share and one
core) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOAD = "511x511x512 blowout-geometry slice loop, 2x2 ppc, shape 2, explicit MG"
ERR_FIELDS = ("Ez", "ExmBy", "jx", "rhomjz")
COPY_BYTES = 256 << 20
# HBM bytes per launch of every kernel family and path, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs scaled by the same counters on a 256 MiB device copy (the commands are in it)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_511x511x512.json")
DTYPE = {"fp64": "f64",
         "fp64_atomic": "f64",
         "fp64_dst": "f64",
         "mixed": "f32 (gather, push increments, deposition) + split-bf16 x3 / f32 accumulate (DST) + f64 "
                  "(multigrid, particle state)"}


def make_workload(n=511, nz=512, ppc=2, nbeam=4_000_000, kind="standard"):
    """The throughput workload (configuration 3), the largest (configuration 5) or the strong
    blowout (configuration 2: 255^2 x 256, 1 x 1 ppc, drive peak 10 n0; one pass of the slice
    loop per step) of hipace2_amd/workloads.py."""
    from hipace2_amd import workloads
    if kind == "large":
        return workloads.config5()
    if kind == "blowout":
        return workloads.config2()
    return workloads.config3(n, nz, ppc, nbeam)


def mg_bytes_per_step(cfg, vcycles):
    """Algorithmic HBM bytes of the multigrid in one step: (18 + 35 k) G 8 per slice with k
    V-cycles on that slice, summed over the slices, (18 nz + 35 K) G 8 with K V-cycles in all:
    18 G words for the solve's set-up (rhs, acf, the guess and the result of two components,
    the first residual and its restriction) and 35 G words per V-cycle (four sweeps down and up
    on every level of a hierarchy whose levels shrink by 4, plus the level-0 residual)."""
    G = cfg.nx * cfg.ny
    return (18 * cfg.nz + 35 * vcycles) * G * 8


def dst_bytes_per_launch(cfg):
    """Algorithmic bytes of one batched row DST launch of the three-field Poisson solve: the x
    passes read and write 3 G words, the fused y pass (forward, eigenvalues, inverse) reads
    3 G + G and writes 3 G; averaged over the three launches of a solve."""
    G = cfg.nx * cfg.ny
    return (6 + 7 + 6) * G * 8 // 3


def bytes_per_launch(kind, cfg, n_plasma, launches=None, vcycles=None):
    """Algorithmic HBM bytes of one launch (each operand read or written once); for the
    multigrid, whose launches differ by level, the step's total over its launch count."""
    G = cfg.nx * cfg.ny
    if kind == "mg":
        return mg_bytes_per_step(cfg, vcycles) / launches if launches and vcycles is not None else None
    if kind == "dst":
        return dst_bytes_per_launch(cfg)
    if kind == "push":
        # read x_prev, y_prev, ux/uy/psi half-step (5 x 8 B) + valid (4 B); write x, y, ux, uy, psi
        # and the five half-step values (10 x 8 B); gather reads Psi, Ez, Bx, By, Bz once
        return n_plasma * (5 * 8 + 4 + 10 * 8) + 5 * G * 8
    if kind == "deposit":
        # read x, y, w, ux, uy, psi (6 x 8 B) + valid; write jx, jy, chi, rhomjz once
        return n_plasma * (6 * 8 + 4) + 4 * G * 8
    if kind == "explicit_deposit":
        # read x, y, w, ux, uy, psi + valid; read Bz, Ez, ExmBy, EypBx; write Sy, Sx
        return n_plasma * (6 * 8 + 4) + 6 * G * 8
    return None


# kernel families: trace kinds of the library (the multigrid's two kinds form one family)
FAMILIES = {"push": ["push"], "deposit": ["deposit"], "explicit_deposit": ["explicit_deposit"],
            "mg": ["mg_smooth", "mg_coarse"], "dst": ["dst"], "beam_deposit": ["beam_deposit"]}


def family_times(traces):
    """(total ms, launches) per kernel family from per-kind traces."""
    out = {}
    for fam, kinds in FAMILIES.items():
        got = [traces[k] for k in kinds if k in traces]
        if got:
            out[fam] = (sum(t[0] for t in got), sum(t[1] for t in got))
    return out


def run_gpu(args, path, capture):
    import torch
    from hipace2_amd.simulation import Simulation

    cfg, beam = make_workload(args.n, args.nz, args.ppc, args.nbeam, kind=args.workload)
    if args.bin_period is not None:
        cfg.bin_period = args.bin_period
    precision = "mixed" if path == "mixed" else "fp64"
    cfg.precision = precision
    cfg.deposition = "atomic" if path in ("mixed", "fp64_atomic", "fp64_dst") else "ordered"
    # the atomic paths reorder the plasma in place by deposition tile (plasmas.reorder_period);
    # the ordered path keeps the index order whose sums it reproduces
    cfg.reorder_period = args.reorder_period if cfg.deposition == "atomic" else 0
    if precision == "mixed":
        cfg.poisson_solver = mixed_poisson(args, cfg)
        cfg.mixed_sums = args.mixed_sums
        cfg.mg_guess = args.mixed_mg_guess
        cfg.mg_guess_iters = args.mixed_mg_guess_iters
    if path == "fp64_dst":
        # FP64 with atomic deposition and the mixed path's DST-corrected multigrid guess
        cfg.mg_guess, cfg.mg_guess_iters = "dst", args.mixed_mg_guess_iters
    sim = Simulation(cfg, beam)
    sim.init_plasma()
    pristine = {k: v.clone() for k, v in sim.plasma.items()}
    kinds = [k for ks in FAMILIES.values() for k in ks]

    def step():
        for k, v in pristine.items():
            sim.plasma[k].copy_(v)
        sim.begin_step(init_plasma=False)
        sim.run_slices()

    # Warm-up steps time every kernel family (an event pair around each launch). In the timed
    # steps only the last one records events, and only around the dominant family's launches:
    # an event pair around each of the multigrid's ~25 launches per slice slows a step by ~17 %,
    # so the others run without events.
    breakdown = {}
    warm_stats = None
    if args.warmup:
        sim.trace(kinds, True)
        sim.trace_reset()
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        warm_stats = sim.stats()
        breakdown = {k: sim.trace_read(k) for k in kinds}
        timed = FAMILIES[dominant_family(family_times(breakdown))]
        sim.trace(kinds, False)
    else:
        timed = kinds
    sim.trace(kinds, False)
    sim.trace_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        if s == args.steps - 1:
            sim.trace(timed, True)
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    traces = {k: sim.trace_read(k) for k in timed}
    stats = sim.stats()
    sim.trace(timed, False)
    caps = None
    if capture:
        caps = {k: sim.capture(k) for k in ERR_FIELDS}
        step()
        torch.cuda.synchronize()
    print(f"{path}: {cfg.nz * args.steps / elapsed:.1f} slices/s", file=sys.stderr, flush=True)
    return dict(cfg=cfg, beam=beam, sim=sim, elapsed=elapsed, traces=traces, stats=stats, breakdown=breakdown,
                warm_stats=warm_stats, caps=caps)


def mixed_poisson(args, cfg):
    """The mixed path's transverse solver: the split-BF16 sine-matrix GEMMs up to n = 511, the
    FFT beyond, where the O(n^3) products lose to the FFT's O(n^2 log n) (both measured)."""
    if args.mixed_poisson != "auto":
        return args.mixed_poisson
    return "bf16x3" if max(cfg.nx, cfg.ny) <= 511 else "fft"


def measured_traffic(kind, path, args):
    """HBM bytes per launch of kernel family `kind` on path `path` (fp64: ordered deposition,
    fp64_atomic, mixed) of this workload, from the committed counter summary, or None when it
    does not cover this family, path and workload."""
    if (args.n, args.nz, args.ppc) != (511, 512, 2) or args.workload != "standard" or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        t = json.load(f)
    e = t.get("families", {}).get(f"{kind}/{path}")
    return None if e is None else int(e["hbm_bytes_per_launch"])


def errors_vs_fp64(mixed, fp64):
    """max |mixed - fp64| / max |fp64| over every slice of one step, per field."""
    out = {}
    for k in ERR_FIELDS:
        a, b = mixed[k], fp64[k]
        den = float(b.abs().max())
        out[k] = float((a - b).abs().max()) / den if den > 0 else float((a - b).abs().max())
    return out


def dominant_family(fam_times):
    """The kernel family that takes the most time."""
    timed = {k: ms for k, (ms, launches) in fam_times.items() if launches}
    return max(timed, key=timed.get)


def family_table(cfg, n_plasma, fam_times, steps, vcycles_per_step):
    """Per family: launches, average launch time, algorithmic bytes per launch, achieved GB/s
    and its fraction of the HBM peak."""
    out = {}
    for fam, (ms, launches) in fam_times.items():
        if not launches:
            continue
        bpl = bytes_per_launch(fam, cfg, n_plasma, launches / steps, vcycles_per_step)
        avg_us = 1e3 * ms / launches
        row = {"total_ms": round(ms, 3), "launches": launches, "avg_us": round(avg_us, 2)}
        if bpl is not None:
            gbs = bpl / (avg_us * 1e-6) / 1e9
            row.update(bytes_per_launch=int(bpl), achieved_GBs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4))
        out[fam] = row
    return out


def run_adjoint(args):
    """Beam-loading gradient: 511 x 511 x 512 slices, 1 x 1 ppc, drive as in the throughput
    workload (1e5 particles), witness of 1e6 particles uniform in zeta on [-6, -4] whose
    current profile is 64 piecewise-linear knots (seeded Uniform[0.5, 1.5]); loss = variance
    of on-axis Ez over the witness slices; checkpoint every 32 slices. Times the plain forward
    step, the checkpointed forward and the backward march, and checks the gradient against
    central differences on `adjoint_fd_knots` knots. Multigrid tolerance 1e-10 for the forward
    solves, so the loss is a smooth function of the knots, and `adjoint_mg_tol` for the
    warm-started adjoint solves."""
    import numpy as np
    import torch
    from hipace2_amd import workloads
    from hipace2_amd.adjoint import on_axis_ez_variance, slice_loop_ez
    from hipace2_amd.simulation import Simulation

    large = args.workload == "large"

    def note(msg):   # progress of the long phases, on stderr
        print(f"adjoint: {msg}", file=sys.stderr, flush=True)

    cfg, bl = workloads.config4(large=large)
    n, nz = cfg.nx, cfg.nz
    P = workloads.LARGE
    sim = Simulation(cfg, bl.beam)
    knots0 = torch.as_tensor(np.random.default_rng(99).uniform(0.5, 1.5, 64), device="cuda")
    wsl = bl.witness_slices()
    # the drive's charge as a factor q on its weights (q = 1 is the nominal drive)
    drive_w = torch.as_tensor(bl.base_w, device="cuda")

    def weights(knots, q=None):
        w = bl.weights(knots)
        return w if q is None else w + (q - 1.0) * drive_w

    def loss_of(ez):
        return on_axis_ez_variance(ez, wsl)

    def plain(knots, q=None):
        w = weights(knots, q)
        sim.beam["w"][: w.numel()].copy_(w)
        ez = sim.capture("Ez")
        sim.begin_step()
        sim.run_slices()
        torch.cuda.synchronize()
        out = float(loss_of(ez))
        sim.clear_captures()
        return out

    plain(knots0)   # warm-up
    t0 = time.perf_counter()
    plain(knots0)
    t_plain = time.perf_counter() - t0
    note(f"plain forward {t_plain:.3f} s")
    # the same forward step at the production multigrid tolerance (hipace.MG_tolerance_rel
    # default 1e-4), the other denominator of the cost ratio
    sim_p = Simulation(dataclasses.replace(cfg, mg_tol_rel=1e-4), bl.beam)
    sim_p.beam["w"][: bl.beam.n].copy_(weights(knots0))
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim_p.begin_step()
        sim_p.run_slices()
        torch.cuda.synchronize()
        t_plain_prod = time.perf_counter() - t0
    del sim_p
    note(f"forward at the production multigrid tolerance {t_plain_prod:.3f} s")

    knots = knots0.clone().requires_grad_(True)
    q = torch.ones((), dtype=torch.float64, device="cuda", requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ez = slice_loop_ez(weights(knots, q), sim, checkpoint_every=32, mg_tol=args.adjoint_mg_tol)
    loss = loss_of(ez)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    grad = knots.grad.detach().cpu().numpy()
    grad_q = float(q.grad)
    note(f"checkpointed forward {t_fwd:.3f} s, backward {t_bwd:.3f} s")

    # the same gradient with the forward (and its recomputation) on the mixed path: FP32
    # particle arithmetic and split-BF16 DST; the vector-Jacobian products stay FP64
    cfg_m = dataclasses.replace(cfg, precision="mixed", poisson_solver=mixed_poisson(args, cfg))
    sim_m = Simulation(cfg_m, bl.beam)
    knots_m = knots0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss_of(slice_loop_ez(bl.weights(knots_m), sim_m, checkpoint_every=32, mg_tol=args.adjoint_mg_tol)).backward()
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    grad_m = knots_m.grad.detach().cpu().numpy()
    mixed_err = float(np.abs(grad_m - grad).max() / np.abs(grad).max())
    del sim_m
    note(f"mixed-path forward + backward {t_mixed:.3f} s")

    picks = np.linspace(0, 63, args.adjoint_fd_knots).round().astype(int) if args.adjoint_fd_knots else []
    fd = []
    for k in picks:
        h = 1e-3 * float(knots0[k])
        e = torch.zeros_like(knots0)
        e[k] = h
        fd.append((plain(knots0 + e) - plain(knots0 - e)) / (2 * h))
        note(f"central difference of knot {k} done")
    fd = np.array(fd)
    g = grad[picks] if len(picks) else np.zeros(0)
    scale = float(np.abs(fd).max()) if fd.size else 0.0
    err = float(np.abs(g - fd).max() / scale) if scale > 0 else None
    # the drive-charge derivative against central differences of q at three step sizes: in a
    # strong blowout the loss is only C1 in the drive charge (order-2 shapes crossing cells),
    # so the difference quotient converges at O(h) and the smallest step is the reference
    fd_q, err_q = {}, None
    if len(picks):
        one = torch.ones((), dtype=torch.float64, device="cuda")
        for hq in (1e-3, 1e-4, 1e-5):
            fd_q[str(hq)] = (plain(knots0, one + hq) - plain(knots0, one - hq)) / (2 * hq)
            note(f"central difference of the drive charge, step {hq:g}, done")
        ref = fd_q[str(1e-5)]
        err_q = abs(grad_q - ref) / abs(ref) if ref else None
    print(f"adjoint: plain {t_plain:.3f} s, checkpointed forward {t_fwd:.3f} s, backward {t_bwd:.3f} s, "
          f"FD max rel err {err}, drive-charge derivative err {err_q}, mixed-forward gradient vs FP64 "
          f"{mixed_err:.3e}", file=sys.stderr, flush=True)
    loading = None if large or args.loading_iters <= 0 else beam_loading(args, cfg, bl, sim, knots0, weights, loss_of)
    plasma = (f"{P['ppc']}x{P['ppc']} ppc with n(r) = n0 (1 + {P['ramp']} r^2), drive 1e7 (peak "
              f"{P['drive_peak']}) + witness 2e6 particles" if large else
              "1x1 ppc, drive 1e5 + witness 1e6 particles")
    return {"workload": f"{n}x{n}x{nz}, {plasma}, 64 knots and the drive charge, loss Var(on-axis Ez) "
                        f"over the witness slices, checkpoint every 32 slices, MG tolerance 1e-10 (adjoint solves {args.adjoint_mg_tol:g}, warm-started), FP64",
            "drive_charge_grad": grad_q, "drive_charge_fd_by_step": fd_q, "drive_charge_rel_err_vs_fd_1e-5": err_q,
            "loss": float(loss.detach()), "plain_forward_s": round(t_plain, 4), "checkpointed_forward_s": round(t_fwd, 4),
            "backward_s": round(t_bwd, 4), "ratio_to_forward": round((t_fwd + t_bwd) / t_plain, 3),
            "plain_forward_production_tol_s": round(t_plain_prod, 4),
            "ratio_to_forward_production_tol": round((t_fwd + t_bwd) / t_plain_prod, 3),
            "fd_knots": [int(k) for k in picks], "fd_max_rel_err": err,
            "mixed_forward_grad_rel_err_vs_fp64": mixed_err, "mixed_forward_plus_backward_s": round(t_mixed, 4),
            "grad_norm": float(np.linalg.norm(grad)), "beam_loading": loading}


def beam_loading(args, cfg, bl, sim, knots0, weights, loss_of):
    """The flatness of the witness's accelerating field reached by gradient descent on the 64
    knots with the adjoint gradient (Adam, `loading_iters` forward + backward passes), against
    SALAME, the reference's own beam-loading algorithm (src/salame/Salame.cpp), which rescales the
    witness slice by slice in one step."""
    import numpy as np
    import torch
    from hipace2_amd.adjoint import slice_loop_ez
    from hipace2_amd.simulation import Simulation

    knots = knots0.clone().requires_grad_(True)
    opt = torch.optim.Adam([knots], lr=0.05)
    history = []
    t0 = time.perf_counter()
    for _ in range(args.loading_iters):
        opt.zero_grad()
        loss = loss_of(slice_loop_ez(weights(knots), sim, checkpoint_every=32, mg_tol=args.adjoint_mg_tol))
        loss.backward()
        history.append(float(loss.detach()))
        opt.step()
        with torch.no_grad():
            knots.clamp_(min=0.0)
    t_opt = time.perf_counter() - t0
    w = weights(knots.detach())
    sim.beam["w"][: w.numel()].copy_(w)
    ez = sim.capture("Ez")
    sim.begin_step()
    sim.run_slices()
    torch.cuda.synchronize()
    loss_adj = float(loss_of(ez))
    sim.clear_captures()

    # SALAME on the witness of the initial knots (its per-slice rescaling replaces the profile)
    cfg_s = dataclasses.replace(cfg, salame=True)
    beam = dataclasses.replace(bl.beam)
    beam.w = weights(knots0).cpu().numpy()
    mask = np.zeros(beam.n, np.int32)
    mask[bl.witness_index] = 1
    beam.salame = mask
    sim_s = Simulation(cfg_s, beam)
    ez = sim_s.capture("Ez")
    sim_s.enable_salame()
    t0 = time.perf_counter()
    sim_s.begin_step()
    sim_s.run_slices()
    torch.cuda.synchronize()
    t_sal = time.perf_counter() - t0
    loss_sal = float(loss_of(ez))
    n_sal = int(np.count_nonzero(sim_s.salame_weight_factor))
    del sim_s
    print(f"beam loading: Var(on-axis Ez) initial {history[0] if history else None:.4e}, adjoint descent "
          f"{loss_adj:.4e} after {args.loading_iters} iterations ({t_opt:.1f} s), SALAME {loss_sal:.4e} "
          f"({t_sal:.2f} s)", file=sys.stderr, flush=True)
    return {"loss_initial": history[0] if history else None, "adjoint_descent_iterations": args.loading_iters,
            "adjoint_descent_loss": loss_adj, "adjoint_descent_s": round(t_opt, 3),
            "adjoint_descent_history": [float(f"{v:.5e}") for v in history],
            "salame_loss": loss_sal, "salame_s": round(t_sal, 3), "salame_slices": n_sal}


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo) and its logical CPUs."""
    name = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count()


def timing_oracle():
    """The oracle built for timing on this host: -O3 -march=native (FMA contraction stays off,
    so its results are those of the committed build), compiled into a temporary directory;
    the committed build when the compiler is not available."""
    import shutil
    import subprocess
    import tempfile
    from oracle import oracle as orc_mod
    src = os.path.join(ROOT, "oracle")
    out = tempfile.mkdtemp(prefix="oracle_native_")
    try:
        for f in os.listdir(src):
            if f.endswith((".cpp", ".hpp")) or f == "Makefile":
                shutil.copy(os.path.join(src, f), out)
        flags = "-O3 -march=native -std=c++17 -fPIC -fopenmp -ffp-contract=off -Wall -Wno-unused-function"
        subprocess.run(["make", "-s", "-C", out, "-j8", f"CXXFLAGS={flags}"], check=True, capture_output=True,
                       timeout=300)
        lib = os.path.join(out, "lib", "liboracle.so")
        orc_mod.use_library(lib)
        return "-O3 -march=native"
    except (OSError, subprocess.SubprocessError) as e:
        print(f"cpu baseline: native oracle build failed ({e}); timing the committed -O3 build", file=sys.stderr)
        return "-O3 (committed build)"


def cpu_baseline(args, cfg, beam, sim):
    """The oracle (the CPU build's algorithms in C++/OpenMP, FP64) on the host cores, started
    from the GPU state at slice s0 (inside the beam): on every core of the box's CPU share and
    on one core."""
    import torch
    from oracle.oracle import OracleSim

    threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
    flags = timing_oracle()
    s0 = min(cfg.nz - 1, int(round((0.44 - cfg.zlo) / cfg.dz)))
    # GPU state entering slice s0
    sim.trace(["push"], False)
    sim.init_plasma()
    sim.begin_step(init_plasma=False)
    sim.run_slices(cfg.nz - 1, s0 + 1)
    torch.cuda.synchronize()
    store, plasma = sim.store_host(), sim.plasma_host()

    def timed(nthreads, budget, max_slices):
        orc = OracleSim(cfg, beam, nthreads=nthreads)
        orc.begin_step()
        orc.set_store(store)
        orc.set_plasma(plasma)
        done = 0
        t0 = time.perf_counter()
        while done < max_slices:
            orc.solve_slice(s0 - done)
            done += 1
            if time.perf_counter() - t0 > budget and done >= 1:
                break
        t = time.perf_counter() - t0
        del orc
        return done, t

    done, t = timed(threads, args.cpu_seconds, args.cpu_max_slices)
    done1, t1 = timed(1, args.cpu_seconds, args.cpu_max_slices)
    model, logical = cpu_model()
    return {"value": round(done / t, 4), "unit": "slices/s", "cores": threads, "kind": "port",
            "one_core": {"value": round(done1 / t1, 4), "unit": "slices/s", "cores": 1,
                         "sample": f"{done1} slices in {t1:.1f} s"},
            "cpu_model": model, "host_logical_cpus": logical, "compiler_flags": flags,
            "sample": f"{done} consecutive slices of the same step starting at slice {s0} (zeta=+0.44, inside "
                      f"the drive beam) from the GPU's state, {t:.1f} s; oracle (C++/OpenMP FP64, the CPU "
                      f"build's algorithms and summation orders) on {threads} threads (the box's CPU share)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=511)
    ap.add_argument("--nz", type=int, default=512)
    ap.add_argument("--ppc", type=int, default=2)
    ap.add_argument("--nbeam", type=int, default=4_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="host threads of the CPU baseline (the CPU share of one GPU on the box)")
    ap.add_argument("--cpu-max-slices", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--bin-period", type=int, default=None,
                    help="slices between tile sorts of the plasma for deposition (0: index order)")
    ap.add_argument("--paths", default="fp64,fp64_atomic,mixed",
                    help="paths to time (comma separated): fp64 (ordered deposition, the headline), fp64_atomic, "
                         "mixed, fp64_dst (atomic deposition with the mixed path's DST-corrected multigrid guess); the "
                         "first is the headline, the others report their error against it")
    ap.add_argument("--mixed-poisson", choices=["auto", "bf16x3", "fft"], default="auto",
                    help="transverse Poisson solver of the mixed path (auto: bf16x3 up to n = 511, fft beyond)")
    ap.add_argument("--mixed-sums", choices=["fp64", "kahan32"], default="fp64",
                    help="window sums of the mixed path's FP32 depositions: FP64 LDS words, or compensated FP32")
    ap.add_argument("--mixed-mg-guess", choices=["previous", "linear", "dst"], default="dst",
                    help="initial guess of the mixed path's multigrid: the previous slice's Bx, By (as the FP64 "
                         "paths and the reference), extrapolated from the two previous slices, or the previous "
                         "slice's corrected by shifted FFT DST solves with chi replaced by its mean")
    ap.add_argument("--mixed-mg-guess-iters", type=int, default=1,
                    help="DST corrections per slice of --mixed-mg-guess dst")
    ap.add_argument("--reorder-period", type=int, default=4,
                    help="plasmas.reorder_period of the atomic paths (fp64_atomic, mixed); 0: never")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the beam-loading gradient measurement")
    ap.add_argument("--adjoint-mg-tol", type=float, default=1e-8,
                    help="relative tolerance of the adjoint multigrid solves (warm-started)")
    ap.add_argument("--adjoint-fd-knots", type=int, default=8,
                    help="knots checked against central differences (two forward steps each)")
    ap.add_argument("--loading-iters", type=int, default=0,
                    help="adjoint gradient-descent iterations of the beam-loading comparison with SALAME (0: skip)")
    ap.add_argument("--workload", choices=["standard", "large", "blowout"], default="standard",
                    help="standard: 511^2 x 512 (the metric's configuration); large: 1023^2 x 1024, 2x2 ppc with "
                         "a density ramp, drive 1e7 + witness 2e6 particles; blowout: 255^2 x 256, 1x1 ppc, "
                         "drive peak 10 n0 (the adjoint measurement stays the 511^2 one)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 256 MiB device to device, the reference dispatch for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the slice loop is not sharded across GPUs "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    paths = [q.strip() for q in args.paths.split(",") if q.strip()]
    for q in paths:
        if q not in ("fp64", "fp64_atomic", "mixed", "fp64_dst"):
            raise SystemExit(f"bench.py: unknown path {q}")
    runs = {q: run_gpu(args, q, capture=len(paths) > 1) for q in paths}
    head = paths[0]
    r = runs[head]
    cfg, beam, sim, elapsed, traces, stats, breakdown = (r[k] for k in ("cfg", "beam", "sim", "elapsed", "traces",
                                                                        "stats", "breakdown"))
    n_plasma = sim.np
    slices = cfg.nz * args.steps
    slices_per_s = slices / elapsed
    # counters are per step (reset by begin_step); the reference counts every plasma particle of
    # a slice as one push (src/particles/pusher/PlasmaParticleAdvance.cpp:70-73)
    pushes_per_s = stats["pushes"] * args.steps / elapsed

    # every family over the warm-up steps (events around every launch), the dominant one over
    # the last timed step (events around its launches only; stats are the last step's)
    warm = family_table(cfg, n_plasma, family_times(breakdown), max(args.warmup, 1),
                        r["warm_stats"]["mg_iters"] if r["warm_stats"] else None) if breakdown else {}
    timed = family_table(cfg, n_plasma, family_times(traces), 1, stats["mg_iters"])
    dom = dominant_family(family_times(traces))
    row = timed[dom]
    roofline = {"bound": "hbm", "achieved": row.get("achieved_GBs"), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": row.get("frac"), "traffic": measured_traffic(dom, head, args),
                "kernel": dom, "bytes_per_launch": row.get("bytes_per_launch"), "avg_launch_us": row["avg_us"],
                "launches": row["launches"], "timed_by": "HIP events around each launch of the family in the last timed step",
                "byte_model": "mg: (18 nz + 35 K) G 8 per step over its launches (K V-cycles)" if dom == "mg" else
                              "each operand read or written once"}
    if warm:
        print(f"kernel families over {args.warmup} warm-up step(s): " + json.dumps(warm), file=sys.stderr)
    print("dominant family over the last timed step: " + json.dumps(timed), file=sys.stderr)

    def summary(q):
        f = runs[q]
        v = cfg.nz * args.steps / f["elapsed"]
        return {"value": round(v, 2), "unit": "slices/s", "ms_per_step": round(1e3 * f["elapsed"] / args.steps, 2),
                "speedup_vs_fp64": round(v / slices_per_s, 3),
                "plasma_pushes_per_s": round(f["stats"]["pushes"] * args.steps / f["elapsed"], 1),
                "mg_vcycles_per_slice": round(f["stats"]["mg_iters"] / cfg.nz, 3), "dtype": DTYPE[q],
                "reorder_period": f["cfg"].reorder_period,
                **({"mixed_sums": f["cfg"].mixed_sums, "mg_guess": f["cfg"].mg_guess,
                    "mg_guess_iters": f["cfg"].mg_guess_iters} if q in ("mixed", "fp64_dst") else {})}

    extra = {}
    for q in paths[1:]:
        extra[q] = summary(q)
        extra[q]["max_rel_err_vs_fp64"] = {k: float(f"{v:.3e}") for k, v in
                                           errors_vs_fp64(runs[q]["caps"], r["caps"]).items()}
    for q in paths:
        runs[q]["caps"] = None
        runs[q]["sim"].clear_captures()

    cpu = None
    if not args.no_cpu:
        sim_cpu = runs["fp64"]["sim"] if "fp64" in runs else sim
        cpu = cpu_baseline(args, cfg, beam, sim_cpu)

    adjoint = None
    if not args.no_adjoint:
        adjoint = run_adjoint(args)

    line = {
        "metric": "slices/s + plasma pushes/s",
        "value": round(slices_per_s, 2),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[head],
        "data": "synthetic (lattice plasma, seeded Philox Gaussian beam)",
        "config": {"workload": WORKLOAD if (args.n, args.nz, args.ppc) == (511, 512, 2) and args.workload == "standard"
                   else f"{cfg.nx}x{cfg.ny}x{cfg.nz} slice loop, {cfg.ppc[0]}x{cfg.ppc[1]} ppc"
                        f"{', density ramp' if cfg.ramp_a else ''}, shape 2, explicit MG",
                   "path": head, "deposition": cfg.deposition, "poisson_solver": cfg.poisson_solver,
                   "nx": cfg.nx, "ny": cfg.ny, "slices": cfg.nz, "plasma_particles_per_slice": n_plasma,
                   "beam_particles": beam.n, "mg_vcycles_per_slice": round(stats["mg_iters"] / cfg.nz, 3),
                   "bin_period": cfg.bin_period},
        "plasma_pushes_per_s": round(pushes_per_s, 1),
        **extra,
        "roofline": roofline,
        "kernel_families": timed,
        "kernel_families_warmup": warm,
        "cpu_baseline": cpu,
        "adjoint": adjoint,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
