"""The oracle against the reference's own golden checksums (CPU/double CI output,
tests/checksum/benchmarks_json/*.json in the reference), and the beam generator against
the golden beam checksums of the same tests."""
import math

import numpy as np
import pytest

from helpers import deck_config, deck_run, golden_cases, run_golden

CASES = sorted(golden_cases().keys())


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_reference_field_checksums(name):
    from oracle.oracle import OracleSim
    case = golden_cases()[name]
    cfg, beam, max_step = deck_run(case)
    orc = OracleSim(cfg, beam, nthreads=4)
    run_golden(orc, cfg, max_step)
    got = orc.checksums()
    # blowout_wake.Serial was produced by an older build and carries its own 2e-2 tolerance;
    # every other golden is matched to 1e-12 (the goldens have 14 significant digits)
    rtol = case["rtol"] if case["rtol"] > 1e-9 else 1e-12
    for field, want in case["fields"].items():
        assert math.isclose(got[field], want, rel_tol=rtol, abs_tol=1e-13), (field, got[field], want)


@pytest.mark.parametrize("name", ["linear_wake.normalized.1Rank", "gaussian_linear_wake.normalized.1Rank",
                                  "beam_in_vacuum.normalized.1Rank"])
def test_beam_generator_matches_reference_beam_checksums(name):
    """fixed_ppc beams: sum |attribute| over all beam particles equals the reference output."""
    case = golden_cases()[name]
    _, beam = deck_config(case)
    g = case["beam"]
    m = beam.valid.astype(bool)
    assert int(m.sum()) == int(g["charge"])
    assert math.isclose(np.abs(beam.w[m]).sum(), g["w"], rel_tol=1e-12)
    assert math.isclose(np.abs(beam.x[m]).sum(), g["x"], rel_tol=1e-12)
    assert math.isclose(np.abs(beam.y[m]).sum(), g["y"], rel_tol=1e-12)
    assert math.isclose(np.abs(beam.z[m]).sum(), g["z"], rel_tol=1e-12)
    assert math.isclose(np.abs(beam.uz[m]).sum(), g["uz"], rel_tol=1e-12)


def beam_checksums(b):
    """sum |attribute| over the valid beam particles (the reference's particle checksums,
    tests/checksum/backend/openpmd_backend.py:47-62)."""
    m = np.asarray(b["valid"]).astype(bool)
    return {k: float(np.abs(np.asarray(b[k])[m]).sum()) for k in ("x", "y", "z", "ux", "uy", "uz", "w")} | \
        {"charge": float(m.sum())}


def test_oracle_reproduces_beam_evolution():
    """beam_evolution.1Rank: 20 beam pushes (10 sub-cycles each) in an external focusing field
    E = (x/2, y/2, 0); the golden beam checksums pin the beam push against the reference."""
    from oracle.oracle import OracleSim
    case = golden_cases()["beam_evolution.1Rank"]
    cfg, beam, max_step = deck_run(case)
    assert max_step == 20 and cfg.dt == 3.0 and beam.external is not None
    orc = OracleSim(cfg, beam, nthreads=4)
    run_golden(orc, cfg, max_step)
    got = beam_checksums(orc.beam())
    for k, want in case["beam"].items():
        if k in got:
            assert math.isclose(got[k], want, rel_tol=1e-12), (k, got[k], want)
