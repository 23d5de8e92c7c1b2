"""Several time steps with a moving beam (hpq_run_step_with_beam + hpq_sort_beam_slices)
against the oracle: fields of the last step within 1e-10 (the deposition sums differ in
order, as for one step) and the beam phase space after every step within 1e-10."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from helpers import rel_err, small_config  # noqa: E402

pytestmark = pytest.mark.gpu


def test_beam_steps_match_oracle():
    from hipace2_amd.inputs import BeamSpec, make_beam
    from hipace2_amd.simulation import Simulation
    from oracle.oracle import OracleSim
    cfg = small_config(n=63, nz=24, order=2, explicit=True)
    spec = BeamSpec(injection_type="fixed_weight", profile="gaussian", density=0.5,
                    position_std=(0.3, 0.3, 1.41), num_particles=20000, u_mean=(0.0, 0.0, 20.0),
                    u_std=(0.5, 0.5, 0.0), seed=5)
    beam = make_beam(cfg, spec)
    sim = Simulation(cfg, beam)
    orc = OracleSim(cfg, beam)
    dt, nsub = 4.0, 10
    for step in range(3):
        if step > 0:
            sim.sort_beam()
            orc.sort_beam()
        if step == 2:
            cg, co = sim.capture("Ez"), orc.capture("Ez")
        sim.run_step_with_beam(dt, nsub)
        orc.begin_step()
        orc.run_step_with_beam(dt, nsub)
        torch.cuda.synchronize()
        ob = orc.beam()
        n = beam.n
        assert np.array_equal(np.asarray(sim.beam_host.off), ob["off"]) or step == 0, step
        assert np.array_equal(sim.beam["valid"][:n].cpu().numpy(), ob["valid"]), step
        for k in ("x", "y", "z", "ux", "uy", "uz"):
            e = rel_err(sim.beam[k][:n].cpu().numpy(), ob[k])
            assert e <= 1e-10, (step, k, e)
    assert rel_err(cg.cpu().numpy(), co) <= 1e-10
