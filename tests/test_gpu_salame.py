"""SALAME on the GPU (hpq_set_salame) against the oracle on the same beams: the per-slice
weight factors and the rescaled weights within 1e-9, the fields within 1e-10 (the deposition
sums differ in order only), and the flattening the algorithm is for."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from helpers import rel_err, salame_case, salame_flatness  # noqa: E402

pytestmark = pytest.mark.gpu


def _pair(**kw):
    from hipace2_amd.simulation import Simulation
    from oracle.oracle import OracleSim
    cfg, beam = salame_case(**{k: v for k, v in kw.items() if k in ("zmin", "zmax")})
    sim, orc = Simulation(cfg, beam), OracleSim(cfg, beam, nthreads=4)
    return cfg, beam, sim, orc


@pytest.mark.parametrize("do_advance,target", [(True, False), (False, False), (True, True)])
def test_salame_matches_oracle(do_advance, target):
    cfg, beam, sim, orc = _pair()
    fn = (lambda z, z0, e0: e0 * (1 + 0.2 * (z0 - z))) if target else None
    eg, jg = sim.capture("Ez"), sim.capture("jz_beam")
    eo, jo = orc.capture("Ez"), orc.capture("jz_beam")
    sim.enable_salame(do_advance=do_advance, target=fn)
    orc.enable_salame(beam.salame, do_advance=do_advance, target=fn)
    sim.run_step()
    orc.run_step()
    torch.cuda.synchronize()
    Wg, Wo = sim.salame_weight_factor, orc.salame_weight_factor
    slices = [int(s) for s in np.nonzero(Wo)[0]]
    assert len(slices) == 6 and np.array_equal(np.nonzero(Wg)[0], np.nonzero(Wo)[0])
    assert rel_err(Wg, Wo) <= 1e-9, (Wg[slices], Wo[slices])
    ob = orc.beam()
    assert np.array_equal(sim.beam["valid"][: beam.n].cpu().numpy(), ob["valid"])
    assert rel_err(sim.beam["w"][: beam.n].cpu().numpy(), ob["w"]) <= 1e-9
    ez = eg.cpu().numpy()
    assert rel_err(ez, eo) <= 1e-10 and rel_err(jg.cpu().numpy(), jo) <= 1e-10
    if do_advance and not target:
        dev = salame_flatness(ez, jg.cpu().numpy(), slices)
        assert max(abs(v) for v in dev.values()) < 1e-4, dev


def test_salame_overloaded_and_off():
    cfg, beam, sim, orc = _pair(zmin=-3.6, zmax=-3.1)
    sim.enable_salame()
    sim.run_step()
    sal = beam.salame.astype(bool)
    valid = sim.beam["valid"][: beam.n].cpu().numpy()
    assert np.all(valid[sal] == 0) and np.array_equal(valid[~sal], beam.valid[~sal])
    # off again: the next step leaves the weights alone
    sim.disable_salame()
    w0 = sim.beam["w"][: beam.n].clone()
    sim.run_step()
    assert torch.equal(sim.beam["w"][: beam.n], w0)


def test_salame_needs_the_explicit_solver():
    from hipace2_amd import _abi
    from hipace2_amd.simulation import Simulation
    cfg, beam = salame_case()
    cfg.explicit_solver = False
    with pytest.raises(_abi.HpqError, match="predictor-corrector"):
        Simulation(cfg, beam)


def test_salame_first_step_of_a_moving_beam_run():
    """run_steps: SALAME in the first time step only (MultiBeam::isSalameNow, step 0), the beam
    pushed after every slice, then a second step with the reshaped weights; against the
    oracle driven the same way."""
    from hipace2_amd.simulation import Simulation
    from oracle.oracle import OracleSim
    cfg, beam = salame_case(nz=32)
    sim, orc = Simulation(cfg, beam), OracleSim(cfg, beam, nthreads=4)
    sim.run_steps(2, dt=2.0, n_subcycles=4)
    orc.enable_salame(beam.salame)
    orc.begin_step()
    orc.run_step_with_beam(2.0, 4)
    orc.disable_salame()
    orc.sort_beam()
    orc.begin_step()
    orc.run_step_with_beam(2.0, 4)
    torch.cuda.synchronize()
    ob = orc.beam()
    assert np.any(sim.salame_weight_factor != 0)
    assert np.array_equal(np.asarray(sim.beam_host.off), ob["off"])
    assert np.array_equal(sim.beam["valid"][: beam.n].cpu().numpy(), ob["valid"])
    for k in ("w", "x", "y", "z", "ux", "uy", "uz"):
        e = rel_err(sim.beam[k][: beam.n].cpu().numpy(), ob[k])
        assert e <= 1e-9, (k, e)
