"""SALAME in the oracle (src/salame/Salame.cpp restated in oracle/sim.cpp). The reference has
no golden for it; it is pinned by what the algorithm is for: with the weights it finds, the
Ez field on the slice behind each SALAME slice, averaged with that slice's beam current,
equals the target (the field where SALAME started, or a user target function) to the
iteration tolerance, while the same beam without SALAME leaves a sloped field."""
import numpy as np
import pytest

from helpers import salame_case, salame_flatness


def _run(n_iter=5, do_advance=True, target=None, salame=True):
    from oracle.oracle import OracleSim
    cfg, beam = salame_case()
    orc = OracleSim(cfg, beam, nthreads=4)
    ez, jz = orc.capture("Ez"), orc.capture("jz_beam")
    if salame:
        orc.enable_salame(beam.salame, n_iter=n_iter, do_advance=do_advance, target=target)
    orc.run_step()
    return cfg, beam, orc, ez, jz


def test_salame_flattens_ez_to_the_initial_field():
    cfg, beam, orc, ez, jz = _run()
    W = orc.salame_weight_factor
    slices = [int(s) for s in np.nonzero(W)[0]]
    lo, hi = int(np.floor((-2.0 - cfg.zlo) / cfg.dz)), int(np.floor((-0.5 - cfg.zlo) / cfg.dz))
    assert slices == list(range(lo, hi)), slices
    dev = salame_flatness(ez, jz, slices)
    assert max(abs(v) for v in dev.values()) < 1e-4, dev
    # converged: the last factors are within the iteration tolerance of 1 (or the last iteration ran)
    assert np.all(np.abs(W[slices] - 1) < 1e-3)
    # the same beam without SALAME: a field that varies along the witness by more than 1 %
    _, _, _, ez0, jz0 = _run(salame=False)
    dev0 = salame_flatness(ez0, jz0, slices)
    assert max(abs(v) for v in dev0.values()) > 1e-2, dev0


def test_salame_follows_a_target_function():
    """hipace.salame_Ez_target(zeta, zeta_initial, Ez_initial) = Ez_initial (1 + 0.2 (zeta_initial - zeta))"""
    fn = lambda z, z0, e0: e0 * (1 + 0.2 * (z0 - z))
    cfg, beam, orc, ez, jz = _run(target=fn)
    slices = [int(s) for s in np.nonzero(orc.salame_weight_factor)[0]]
    s0 = max(slices)
    zeta = lambda s: s * cfg.dz + 0.5 * (cfg.zlo + cfg.zhi - cfg.dz * (cfg.nz - 1))
    e0 = {s: float((jz[s] * ez[s0]).sum() / jz[s].sum()) for s in slices}
    dev = salame_flatness(ez, jz, slices, target=lambda s: fn(zeta(s - 1), zeta(s0), e0[s]))
    assert max(abs(v) for v in dev.values()) < 1e-4, dev


def test_salame_without_plasma_advance_is_an_approximation():
    """hipace.salame_do_advance = 0: the SALAME-only currents come from chi and B
    (SalameGetJxJyFromBxBy), an approximation of the plasma response."""
    cfg, beam, orc, ez, jz = _run(do_advance=False)
    slices = [int(s) for s in np.nonzero(orc.salame_weight_factor)[0]]
    dev = salame_flatness(ez, jz, slices)
    assert max(abs(v) for v in dev.values()) < 2e-2, dev


def test_salame_overloaded_beam_is_removed():
    """A witness in the decelerating phase asks for a negative weight: SALAME sets it to zero
    (SalameGetW :136-143) and the particles of that slice are invalidated."""
    from oracle.oracle import OracleSim
    cfg, beam = salame_case(zmin=-3.6, zmax=-3.1)
    orc = OracleSim(cfg, beam, nthreads=4)
    orc.enable_salame(beam.salame)
    orc.run_step()
    b = orc.beam()
    sal = beam.salame.astype(bool)
    assert np.all(b["valid"][sal] == 0)
    assert np.array_equal(b["valid"][~sal], beam.valid[~sal])
