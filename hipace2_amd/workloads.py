"""The five benchmark workloads of the per-slice loop, defined once for the benchmark and the
GPU tests. All in normalised units with Dirichlet fields; seeds fixed.

1. Plain FP64 linear wake: 63 x 63 over [-6, 6]^2, 100 slices over zeta in [-8, 4], uniform
   plasma n0 = 1 at 1 x 1 ppc, shape order 1; Gaussian drive of peak 0.01 n0, sigma_r 0.3,
   sigma_z 1.41, 1e5 particles, gamma 1000; one step, predictor-corrector Bx/By solver.
2. Plain FP64 blowout: 255 x 255 over [-8, 8]^2, 256 slices over zeta in [-6, 6] (the reference's
   examples/blowout_wake/inputs_normalized), 1 x 1 ppc, shape order 2; drive of peak 10 n0,
   sigma_r 0.1, sigma_z 1.41, 1e6 particles, gamma 2000; explicit solver with multigrid; ten
   steps of dt = 10 (a betatron period of the drive is 2 pi sqrt(2 gamma) ~ 400) with ten beam
   sub-cycles.
3. Blowout geometry at 511 x 511 over [-8, 8]^2, 512 slices over [-8, 4], 2 x 2 ppc, shape 2; drive
   as in 1 with 4e6 particles; explicit solver (the throughput workload).
4. Beam loading: 511 x 511, 512 slices, 1 x 1 ppc; drive as in 1 (1e5 particles); witness of 1e6
   particles uniform in zeta on [-6, -4] whose current profile is 64 piecewise-linear knots;
   loss = variance of on-axis Ez over the witness slices; checkpoints every 32 slices.
5. Large: 1023 x 1023 over [-10, 10]^2, 1024 slices, 2 x 2 ppc with n(r) = n0 (1 + 0.1 r^2), drive of
   1e7 particles at peak 5 n0, witness of 2e6.
"""
from __future__ import annotations

from .inputs import BeamSpec, SliceConfig, make_beam, merge_beams

# workload 5's sizes
LARGE = dict(n=1023, nz=1024, half_width=10.0, ppc=2, ramp=0.1, drive_peak=5.0, drive_n=10_000_000,
             witness_n=2_000_000, witness_peak=1.0)
CONFIG2_DT, CONFIG2_STEPS, CONFIG2_SUBCYCLES = 10.0, 10, 10


def _drive(cfg, peak, sigma_r, n, uz, seed, sigma_z=1.41):
    return make_beam(cfg, BeamSpec(injection_type="fixed_weight", profile="gaussian", density=peak,
                                   position_std=(sigma_r, sigma_r, sigma_z), num_particles=n,
                                   u_mean=(0.0, 0.0, uz), seed=seed))


def config1():
    cfg = SliceConfig(nx=63, ny=63, nz=100, xlo=-6.0, xhi=6.0, ylo=-6.0, yhi=6.0, zlo=-8.0, zhi=4.0,
                      depos_order=1, derivative_type=2, explicit_solver=False, ppc=(1, 1))
    return cfg, _drive(cfg, 0.01, 0.3, 100_000, 1000.0, seed=1001)


def config2():
    """(cfg, beam); the step is CONFIG2_DT with CONFIG2_SUBCYCLES beam sub-cycles, CONFIG2_STEPS
    steps."""
    cfg = SliceConfig(nx=255, ny=255, nz=256, xlo=-8.0, xhi=8.0, ylo=-8.0, yhi=8.0, zlo=-6.0, zhi=6.0,
                      depos_order=2, derivative_type=2, explicit_solver=True, ppc=(1, 1))
    cfg.dt, cfg.n_subcycles = CONFIG2_DT, CONFIG2_SUBCYCLES
    return cfg, _drive(cfg, 10.0, 0.1, 1_000_000, 2000.0, seed=1002)


def config3(n=511, nz=512, ppc=2, nbeam=4_000_000):
    cfg = SliceConfig(nx=n, ny=n, nz=nz, xlo=-8.0, xhi=8.0, ylo=-8.0, yhi=8.0, zlo=-8.0, zhi=4.0,
                      depos_order=2, derivative_type=2, explicit_solver=True, ppc=(ppc, ppc))
    return cfg, _drive(cfg, 0.01, 0.3, nbeam, 1000.0, seed=1234)


def config4(large=False):
    """(cfg, BeamLoading) of the beam-loading gradient (large: at workload 5's sizes). The
    forward multigrid tolerance is 1e-10 so that the loss is smooth in the knots."""
    from .adjoint import BeamLoading, witness_beam
    if large:
        P = LARGE
        n, nz, L = P["n"], P["nz"], P["half_width"]
        cfg = SliceConfig(nx=n, ny=n, nz=nz, xlo=-L, xhi=L, ylo=-L, yhi=L, zlo=-8.0, zhi=4.0, depos_order=2,
                          derivative_type=2, explicit_solver=True, ppc=(P["ppc"], P["ppc"]), ramp_a=P["ramp"])
    else:
        cfg = SliceConfig(nx=511, ny=511, nz=512, xlo=-8.0, xhi=8.0, ylo=-8.0, yhi=8.0, zlo=-8.0, zhi=4.0,
                          depos_order=2, derivative_type=2, explicit_solver=True, ppc=(1, 1))
    cfg.do_beam_jx_jy = False
    cfg.mg_tol_rel = 1e-10
    drive = _drive(cfg, LARGE["drive_peak"] if large else 0.01, 0.3, LARGE["drive_n"] if large else 100_000,
                   1000.0, seed=1234)
    wit = witness_beam(cfg, LARGE["witness_n"] if large else 1_000_000, -6.0, -4.0, 0.2,
                       LARGE["witness_peak"] if large else 0.005, 1000.0, seed=4321)
    return cfg, BeamLoading(cfg, drive, wit, -6.0, -4.0)


def config5():
    P = LARGE
    L = P["half_width"]
    cfg = SliceConfig(nx=P["n"], ny=P["n"], nz=P["nz"], xlo=-L, xhi=L, ylo=-L, yhi=L, zlo=-8.0, zhi=4.0,
                      depos_order=2, derivative_type=2, explicit_solver=True, ppc=(P["ppc"], P["ppc"]),
                      ramp_a=P["ramp"])
    from .adjoint import witness_beam
    drive = _drive(cfg, P["drive_peak"], 0.3, P["drive_n"], 1000.0, seed=1234)
    wit = witness_beam(cfg, P["witness_n"], -6.0, -4.0, 0.2, P["witness_peak"], 1000.0, seed=4321)
    return cfg, merge_beams(cfg, [drive, wit])
