"""Throughput of the per-slice quasi-static loop on one MI355X.

Workload (one step = one pass of the slice loop over all ζ-slices, inputs resident in HBM):
  511 x 511 transverse grid over [-8, 8]^2, 512 slices over ζ in [-8, 4], uniform plasma
  n0 = 1 at 2 x 2 particles per cell (1 044 484 plasma particles per slice), shape order 2,
  explicit Bx/By solver with multigrid; Gaussian electron drive beam with peak density
  0.01 n0, sigma_r = 0.3, sigma_z = 1.41, 4e6 macro-particles (seeded Philox), gamma = 1000.

Two arithmetic paths are timed on the same workload: the FP64 plain path (the drop-in)
and the mixed path (FP32 gather, push increments and deposition, the FP32 contributions
summed in FP64 LDS words; DST-I as split-BF16 sine-matrix GEMMs on the matrix cores; FP64 multigrid
and particle state). One more untimed step of each records
Ez, Ex-By, jx and rho-jz on every slice, and the mixed path's norm-wise relative error
against the FP64 path is reported.

Prints one JSON line: slices/s of the mixed path (value) beside the FP64 path's, plasma
particle pushes/s, the roofline of the dominant kernel (HIP-event timed inside the timed
region) and the oracle timed on the host cores on a bounded sample of the same workload.
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
# HBM bytes per launch of the dominant kernels, from rocprofv3 FETCH_SIZE and WRITE_SIZE runs
# scaled by the same counters on a 256 MiB device copy (see the summary beside it)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_511x511x512.json")
DTYPE = {"fp64": "f64",
         "mixed": "f32 (gather, push increments, deposition) + split-bf16 x3 / f32 accumulate (DST) + f64 "
                  "(multigrid, particle state)"}


# The large workload: 1023 x 1023 over [-10, 10]^2, 1024 slices, 2 x 2 ppc with the transverse
# density ramp n(r) = n0 (1 + 0.1 r^2), drive of 1e7 particles at peak 5 n0, witness of 2e6
LARGE = dict(n=1023, nz=1024, half_width=10.0, ppc=2, ramp=0.1, drive_peak=5.0, drive_n=10_000_000,
             witness_n=2_000_000, witness_peak=1.0)


def make_workload(n=511, nz=512, ppc=2, nbeam=4_000_000, large=False):
    from hipace2_amd.inputs import BeamSpec, SliceConfig, make_beam, merge_beams
    if large:
        from hipace2_amd.adjoint import witness_beam
        P = LARGE
        L = P["half_width"]
        cfg = SliceConfig(nx=P["n"], ny=P["n"], nz=P["nz"], xlo=-L, xhi=L, ylo=-L, yhi=L, zlo=-8.0, zhi=4.0,
                          depos_order=2, derivative_type=2, explicit_solver=True, ppc=(P["ppc"], P["ppc"]),
                          ramp_a=P["ramp"])
        drive = make_beam(cfg, BeamSpec(injection_type="fixed_weight", profile="gaussian", density=P["drive_peak"],
                                        position_std=(0.3, 0.3, 1.41), num_particles=P["drive_n"],
                                        u_mean=(0.0, 0.0, 1000.0), seed=1234))
        wit = witness_beam(cfg, P["witness_n"], -6.0, -4.0, 0.2, P["witness_peak"], 1000.0, seed=4321)
        return cfg, merge_beams(cfg, [drive, wit])
    cfg = SliceConfig(nx=n, ny=n, nz=nz, xlo=-8.0, xhi=8.0, ylo=-8.0, yhi=8.0, zlo=-8.0, zhi=4.0,
                      depos_order=2, derivative_type=2, explicit_solver=True, ppc=(ppc, ppc))
    spec = BeamSpec(injection_type="fixed_weight", profile="gaussian", density=0.01,
                    position_std=(0.3, 0.3, 1.41), num_particles=nbeam, u_mean=(0.0, 0.0, 1000.0), seed=1234)
    return cfg, make_beam(cfg, spec)


def bytes_per_launch(kind, cfg, n_plasma):
    """Algorithmic HBM bytes of one launch (each operand read or written once)."""
    G = cfg.nx * cfg.ny
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


def run_gpu(args, precision, capture):
    import torch
    from hipace2_amd.simulation import Simulation

    cfg, beam = make_workload(args.n, args.nz, args.ppc, args.nbeam, large=args.workload == "large")
    if args.bin_period is not None:
        cfg.bin_period = args.bin_period
    cfg.precision = precision
    if precision == "mixed":
        cfg.poisson_solver = args.mixed_poisson
    sim = Simulation(cfg, beam)
    sim.init_plasma()
    pristine = {k: v.clone() for k, v in sim.plasma.items()}
    kinds = ["push", "deposit", "explicit_deposit", "mg_smooth", "mg_coarse", "dst", "beam_deposit"]

    def step():
        for k, v in pristine.items():
            sim.plasma[k].copy_(v)
        sim.begin_step(init_plasma=False)
        sim.run_slices()

    # Warm-up steps time every kernel family (an event pair around each launch); the
    # timed steps keep events only around the dominant family, so the timed region is not
    # slowed by tens of event records per slice.
    breakdown = {}
    if args.warmup:
        sim.trace(kinds, True)
        sim.trace_reset()
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        breakdown = {k: sim.trace_read(k) for k in kinds}
        timed = [dominant_kind(cfg, sim.np, breakdown)]
        sim.trace(kinds, False)
    else:
        timed = kinds
    sim.trace(timed, True)
    sim.trace_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
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
    print(f"{precision}: {cfg.nz * args.steps / elapsed:.1f} slices/s", file=sys.stderr, flush=True)
    return dict(cfg=cfg, beam=beam, sim=sim, elapsed=elapsed, traces=traces, stats=stats, breakdown=breakdown,
                caps=caps)


def measured_traffic(kind, precision, args):
    """HBM bytes per launch of kernel family `kind` on this workload, from the committed
    counter summary, or None when it does not cover this kernel and workload."""
    if (args.n, args.nz, args.ppc) != (511, 512, 2) or args.workload != "standard" or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        t = json.load(f)
    e = t.get("kernels", {}).get(f"{kind}/{precision}")
    return None if e is None else int(e["hbm_bytes_per_launch"])


def errors_vs_fp64(mixed, fp64):
    """max |mixed - fp64| / max |fp64| over every slice of one step, per field."""
    out = {}
    for k in ERR_FIELDS:
        a, b = mixed[k], fp64[k]
        den = float(b.abs().max())
        out[k] = float((a - b).abs().max()) / den if den > 0 else float((a - b).abs().max())
    return out


def dominant_kind(cfg, n_plasma, traces):
    """The kernel family with an algorithmic byte count that takes the most time."""
    bounded = {k: ms for k, (ms, launches) in traces.items()
               if launches and bytes_per_launch(k, cfg, n_plasma) is not None}
    return max(bounded, key=bounded.get)


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
    from hipace2_amd.adjoint import BeamLoading, on_axis_ez_variance, slice_loop_ez, witness_beam
    from hipace2_amd.inputs import BeamSpec, SliceConfig, make_beam
    from hipace2_amd.simulation import Simulation

    large = args.workload == "large"
    if large:
        P = LARGE
        n, nz, L = P["n"], P["nz"], P["half_width"]
        cfg = SliceConfig(nx=n, ny=n, nz=nz, xlo=-L, xhi=L, ylo=-L, yhi=L, zlo=-8.0, zhi=4.0, depos_order=2,
                          derivative_type=2, explicit_solver=True, ppc=(P["ppc"], P["ppc"]), ramp_a=P["ramp"])
    else:
        n, nz = args.n, args.nz
        cfg = SliceConfig(nx=n, ny=n, nz=nz, xlo=-8.0, xhi=8.0, ylo=-8.0, yhi=8.0, zlo=-8.0, zhi=4.0,
                          depos_order=2, derivative_type=2, explicit_solver=True, ppc=(1, 1))
    cfg.do_beam_jx_jy = False
    cfg.mg_tol_rel = 1e-10
    drive = make_beam(cfg, BeamSpec(injection_type="fixed_weight", profile="gaussian",
                                    density=LARGE["drive_peak"] if large else 0.01,
                                    position_std=(0.3, 0.3, 1.41),
                                    num_particles=LARGE["drive_n"] if large else 100_000,
                                    u_mean=(0.0, 0.0, 1000.0), seed=1234))
    wit = witness_beam(cfg, LARGE["witness_n"] if large else 1_000_000, -6.0, -4.0, 0.2,
                       LARGE["witness_peak"] if large else 0.005, 1000.0, seed=4321)
    bl = BeamLoading(cfg, drive, wit, -6.0, -4.0)
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

    # the same gradient with the forward (and its recomputation) on the mixed path: FP32
    # particle arithmetic and split-BF16 DST; the vector-Jacobian products stay FP64
    cfg_m = dataclasses.replace(cfg, precision="mixed", poisson_solver="bf16x3")
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

    picks = np.linspace(0, 63, args.adjoint_fd_knots).round().astype(int) if args.adjoint_fd_knots else []
    fd = []
    for k in picks:
        h = 1e-3 * float(knots0[k])
        e = torch.zeros_like(knots0)
        e[k] = h
        fd.append((plain(knots0 + e) - plain(knots0 - e)) / (2 * h))
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
        ref = fd_q[str(1e-5)]
        err_q = abs(grad_q - ref) / abs(ref) if ref else None
    print(f"adjoint: plain {t_plain:.3f} s, checkpointed forward {t_fwd:.3f} s, backward {t_bwd:.3f} s, "
          f"FD max rel err {err}, drive-charge derivative err {err_q}, mixed-forward gradient vs FP64 "
          f"{mixed_err:.3e}", file=sys.stderr, flush=True)
    loading = None if large or args.loading_iters <= 0 else beam_loading(args, cfg, bl, sim, knots0, weights, loss_of)
    plasma = (f"{LARGE['ppc']}x{LARGE['ppc']} ppc with n(r) = n0 (1 + {LARGE['ramp']} r^2), drive 1e7 (peak "
              f"{LARGE['drive_peak']}) + witness 2e6 particles" if large else
              "1x1 ppc, drive 1e5 + witness 1e6 particles")
    return {"workload": f"{n}x{n}x{nz}, {plasma}, 64 knots and the drive charge, loss Var(on-axis Ez) "
                        f"over the witness slices, checkpoint every 32 slices, MG tolerance 1e-10 (adjoint solves {args.adjoint_mg_tol:g}, warm-started), FP64",
            "drive_charge_grad": grad_q, "drive_charge_fd_by_step": fd_q, "drive_charge_rel_err_vs_fd_1e-5": err_q,
            "loss": float(loss.detach()), "plain_forward_s": round(t_plain, 4), "checkpointed_forward_s": round(t_fwd, 4),
            "backward_s": round(t_bwd, 4), "ratio_to_forward": round((t_fwd + t_bwd) / t_plain, 3),
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


def cpu_baseline(args, cfg, beam, sim):
    """The oracle on the host cores, started from the GPU state at slice s0 (inside the beam)."""
    import numpy as np
    import torch
    from oracle.oracle import OracleSim
    from hipace2_amd.simulation import PLASMA_REALS

    threads = max(1, min(16, os.cpu_count() or 1))
    s0 = min(cfg.nz - 1, int(round((0.44 - cfg.zlo) / cfg.dz)))
    # GPU state entering slice s0
    sim.trace(["push"], False)
    sim.init_plasma()
    sim.begin_step(init_plasma=False)
    sim.run_slices(cfg.nz - 1, s0 + 1)
    torch.cuda.synchronize()
    orc = OracleSim(cfg, beam, nthreads=threads)
    orc.begin_step()
    orc.set_store(sim.store_host())
    p = sim.plasma_host()
    orc.set_plasma(p)
    done = 0
    t0 = time.perf_counter()
    budget = args.cpu_seconds
    while done < args.cpu_max_slices:
        orc.solve_slice(s0 - done)
        done += 1
        if time.perf_counter() - t0 > budget and done >= 2:
            break
    t = time.perf_counter() - t0
    return {"value": done / t, "unit": "slices/s", "cores": threads, "kind": "port",
            "sample": f"{done} consecutive slices of the same step starting at slice {s0} (zeta=+0.44, inside "
                      f"the drive beam) from the GPU's state, oracle (C++/OpenMP FP64) on {threads} host threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=511)
    ap.add_argument("--nz", type=int, default=512)
    ap.add_argument("--ppc", type=int, default=2)
    ap.add_argument("--nbeam", type=int, default=4_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-max-slices", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--bin-period", type=int, default=None,
                    help="slices between tile sorts of the plasma for deposition (0: index order)")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both",
                    help="arithmetic path(s) to time; 'both' also reports the mixed path's error")
    ap.add_argument("--mixed-poisson", choices=["bf16x3", "fft"], default="bf16x3",
                    help="transverse Poisson solver of the mixed path")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the beam-loading gradient measurement")
    ap.add_argument("--adjoint-mg-tol", type=float, default=1e-8,
                    help="relative tolerance of the adjoint multigrid solves (warm-started)")
    ap.add_argument("--adjoint-fd-knots", type=int, default=8,
                    help="knots checked against central differences (two forward steps each)")
    ap.add_argument("--loading-iters", type=int, default=0,
                    help="adjoint gradient-descent iterations of the beam-loading comparison with SALAME (0: skip)")
    ap.add_argument("--workload", choices=["standard", "large"], default="standard",
                    help="standard: 511^2 x 512 (the metric's configuration); large: 1023^2 x 1024, 2x2 ppc with "
                         "a density ramp, drive 1e7 + witness 2e6 particles")
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
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    runs = {p: run_gpu(args, p, capture=len(precs) == 2) for p in precs}
    head = precs[-1]
    r = runs[head]
    cfg, beam, sim, elapsed, traces, stats, breakdown = (r[k] for k in ("cfg", "beam", "sim", "elapsed", "traces",
                                                                        "stats", "breakdown"))
    n_plasma = sim.np
    slices = cfg.nz * args.steps
    slices_per_s = slices / elapsed
    # counters are per step (reset by begin_step); the reference counts every plasma particle of
    # a slice as one push (src/particles/pusher/PlasmaParticleAdvance.cpp:70-73)
    pushes_per_s = stats["pushes"] * args.steps / elapsed

    def table(tr):
        return {k: {"total_ms": round(ms, 3), "launches": launches,
                    "avg_us": round(1e3 * ms / launches, 2) if launches else 0.0} for k, (ms, launches) in tr.items()}

    rows = table(traces)
    dom = dominant_kind(cfg, n_plasma, traces)
    bpl = bytes_per_launch(dom, cfg, n_plasma)
    avg_s = rows[dom]["avg_us"] * 1e-6
    achieved = bpl / avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, head, args),
                "kernel": dom, "bytes_per_launch": bpl, "avg_launch_us": round(rows[dom]["avg_us"], 2)}
    if breakdown:
        print(f"kernel times over {args.warmup} warm-up step(s): " + json.dumps(table(breakdown)), file=sys.stderr)
    print("dominant kernel over the timed region: " + json.dumps(rows), file=sys.stderr)

    extra = {}
    if len(precs) == 2:
        f = runs["fp64"]
        f_slices = cfg.nz * args.steps / f["elapsed"]
        extra["fp64"] = {"value": round(f_slices, 2), "unit": "slices/s",
                         "ms_per_step": round(1e3 * f["elapsed"] / args.steps, 2),
                         "plasma_pushes_per_s": round(f["stats"]["pushes"] * args.steps / f["elapsed"], 1),
                         "mg_vcycles_per_slice": round(f["stats"]["mg_iters"] / cfg.nz, 3)}
        extra["speedup_vs_fp64"] = round(slices_per_s / f_slices, 3)
        extra["max_rel_err_vs_fp64"] = {k: float(f"{v:.3e}") for k, v in
                                        errors_vs_fp64(r["caps"], f["caps"]).items()}
        for p in precs:
            runs[p]["caps"] = None
            runs[p]["sim"].clear_captures()

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
                   "precision": head, "poisson_solver": cfg.poisson_solver,
                   "nx": cfg.nx, "ny": cfg.ny, "slices": cfg.nz, "plasma_particles_per_slice": n_plasma,
                   "beam_particles": beam.n, "mg_vcycles_per_slice": round(stats["mg_iters"] / cfg.nz, 3),
                   "bin_period": cfg.bin_period},
        "plasma_pushes_per_s": round(pushes_per_s, 1),
        **extra,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "adjoint": adjoint,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
