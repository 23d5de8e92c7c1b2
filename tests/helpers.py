"""Shared helpers for the parity tests: configurations, seeded random states, norms."""
import json
import math
import os

import numpy as np

from hipace2_amd.inputs import SliceConfig, parse_deck, make_beam, merge_beams

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "reference_checksums.json")

PLASMA_REALS = ["x", "y", "w", "ux", "uy", "psi", "x_prev", "y_prev", "ux_half_step", "uy_half_step",
                "psi_half_step"]


def golden_cases():
    with open(GOLDEN) as f:
        return json.load(f)


def deck_config(case):
    """(cfg, beam) for one golden case of tests/golden/reference_checksums.json."""
    cfg, beam, _ = deck_run(case)
    return cfg, beam


def deck_run(case):
    """(cfg, beam, max_step) for one golden case."""
    text = "\n".join(f"{k} = {v}" for k, v in case["deck"].items())
    cfg, beams, max_step = parse_deck(text, case["overrides"])
    beam = merge_beams(cfg, [make_beam(cfg, b) for b in beams])
    return cfg, beam, max_step


def run_golden(sim, cfg, max_step):
    """Evolve's steps 0..max_step as the reference's test ran them; the output (checksums) is
    that of the last step, whose fields do not depend on that step's beam push. With dt = 0
    every step repeats the first, so one step stands for all. `sim` is a Simulation or an
    OracleSim; afterwards its beam holds the state the last step's output was taken from."""
    if cfg.dt == 0.0 or max_step == 0:
        sim.run_step()
        return
    for step in range(max_step + 1):
        if step > 0:
            sim.sort_beam()
        if step == max_step:
            sim.run_step()
        elif hasattr(sim, "ctx"):
            sim.run_step_with_beam(cfg.dt, cfg.n_subcycles, cfg.do_z_push)
        else:
            sim.begin_step()
            sim.run_step_with_beam(cfg.dt, cfg.n_subcycles, cfg.do_z_push)


def small_config(n=63, nz=16, order=2, dtype=2, explicit=True, L=6.0):
    return SliceConfig(nx=n, ny=n, nz=nz, xlo=-L, xhi=L, ylo=-L, yhi=L, zlo=-8.0, zhi=4.0,
                       depos_order=order, derivative_type=dtype, explicit_solver=explicit)


def random_plasma(cfg, n, seed=0, spread=0.2):
    """Seeded plasma state with every particle strictly inside the domain."""
    rng = np.random.default_rng(seed)
    d = {}
    m = 0.5 * cfg.dx
    d["x"] = rng.uniform(cfg.xlo + m, cfg.xhi - m, n)
    d["y"] = rng.uniform(cfg.ylo + m, cfg.yhi - m, n)
    d["w"] = rng.uniform(0.5, 1.5, n)
    d["ux"] = rng.normal(0, spread, n)
    d["uy"] = rng.normal(0, spread, n)
    d["psi"] = 1.0 + rng.uniform(-0.3, 0.3, n)
    d["x_prev"] = np.clip(d["x"] + rng.normal(0, 0.1 * cfg.dx, n), cfg.xlo + m, cfg.xhi - m)
    d["y_prev"] = np.clip(d["y"] + rng.normal(0, 0.1 * cfg.dy, n), cfg.ylo + m, cfg.yhi - m)
    d["ux_half_step"] = d["ux"] + rng.normal(0, 0.01, n)
    d["uy_half_step"] = d["uy"] + rng.normal(0, 0.01, n)
    d["psi_half_step"] = d["psi"] + rng.uniform(-0.01, 0.01, n)
    d["valid"] = (rng.uniform(0, 1, n) > 0.05).astype(np.int32)
    return d


def random_store(ncomp, ldy, ldx, seed=1, scale=1.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(-scale, scale, (ncomp, ldy, ldx))


def rel_err(a, b):
    """max |a - b| / max |b| (norm-wise relative error, 0 if both vanish)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if a.size else 0.0
    if den == 0.0:
        return num
    return num / den


def salame_case(nz=48, do_salame=True, zmin=-2.0, zmax=-0.5):
    """(cfg, beam): a drive beam and, behind it in the accelerating phase of its wake, a
    flat-top witness that SALAME reshapes (<beam>.do_salame)."""
    from hipace2_amd.inputs import BeamSpec
    cfg = SliceConfig(nx=63, ny=63, nz=nz, xlo=-6, xhi=6, ylo=-6, yhi=6, zlo=-8.0, zhi=4.0)
    cfg.salame = do_salame
    drive = BeamSpec(injection_type="fixed_weight", profile="gaussian", density=2.0, position_mean=(0, 0, 2.0),
                     position_std=(0.3, 0.3, 0.6), num_particles=20000, u_mean=(0, 0, 1000.0), seed=1)
    wit = BeamSpec(injection_type="fixed_ppc", profile="flattop", density=0.5, zmin=zmin, zmax=zmax, radius=0.4,
                   ppc=(2, 2, 1), u_mean=(0, 0, 1000.0), do_salame=do_salame)
    return cfg, merge_beams(cfg, [make_beam(cfg, drive), make_beam(cfg, wit)])


def salame_flatness(ez, jz, slices, target=None):
    """Relative deviation, per SALAME slice s, of the jz(s)-weighted mean of Ez on the next
    slice (s - 1) from the target: the jz(s)-weighted mean of the first SALAME slice's Ez
    plane, or target(s) when given."""
    s0 = max(slices)
    out = {}
    for s in sorted(slices, reverse=True):
        if s == 0:
            continue
        j = jz[s]
        got = float((j * ez[s - 1]).sum() / j.sum())
        want = float((j * ez[s0]).sum() / j.sum()) if target is None else target(s)
        out[s] = (got - want) / abs(want)
    return out
