"""Host-side input handling: the external-field expressions of the reference's input decks
(beams.external_E/B(x,y,z,t), src/particles/beam/BeamParticleContainer.cpp:72-87) as affine
coefficients, and the deck keys of the beam push."""
import numpy as np
import pytest

from hipace2_amd.inputs import external_fields, parse_affine, parse_deck


@pytest.mark.parametrize("expr,want", [
    (".5*x", [0, 0.5, 0, 0]),
    ("0.", [0, 0, 0, 0]),
    ("-2*(y - 1)/4", [0.5, 0, -0.5, 0]),
    ("2^3*z + x/2 - 3", [-3, 0.5, 0, 8]),
    ('"1e-3*(x+y)"', [0, 1e-3, 1e-3, 0]),
])
def test_affine_coefficients(expr, want):
    np.testing.assert_array_equal(parse_affine(expr), np.array(want, dtype=np.float64))


@pytest.mark.parametrize("expr", ["x*y", "x^2", "1/x", "t", "sin(x)", "__import__('os')"])
def test_non_affine_rejected(expr):
    with pytest.raises(ValueError):
        parse_affine(expr)


def test_external_fields_layout_and_deck():
    e = external_fields([".5*x", ".5*y", "0."], None)
    assert e.shape == (6, 4) and e[0, 1] == 0.5 and e[1, 2] == 0.5 and not e[2:].any()
    assert external_fields(None, None) is None
    deck = "\n".join(["amr.n_cell = 8 8 4", "geometry.prob_lo = -1 -1 -1", "geometry.prob_hi = 1 1 1",
                      "hipace.dt = 3.", "beams.names = beam", "beam.injection_type = fixed_ppc",
                      "beam.density = 1", "beam.n_subcycles = 4",
                      "beams.external_E(x,y,z,t) = .5*x .5*y 0.", "beams.external_B(x,y,z,t) = 0 0 z"])
    cfg, beams, _ = parse_deck(deck)
    assert cfg.dt == 3.0 and cfg.n_subcycles == 4 and cfg.do_z_push
    ext = beams[0].external
    assert ext[0, 1] == 0.5 and ext[1, 2] == 0.5 and ext[5, 3] == 1.0 and ext[3:5].sum() == 0
