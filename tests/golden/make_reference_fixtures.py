"""Writes tests/golden/reference_checksums.json from the reference's own test data.

Inputs (read as text from a checkout of the reference at REF, default /root/reference):
  examples/<case>/inputs_normalized               input decks (key = value data)
  tests/<test>.sh                                  only the override lines are transcribed below
  tests/checksum/benchmarks_json/<test>.json       golden checksums of the CPU/double CI

The fixture holds, per test: the deck's key/value pairs with the test script's
overrides applied, the checksum tolerance the test script uses, and the golden
field checksums (sum of |F| over the xyz output of the last step,
tests/checksum/backend/openpmd_backend.py:40-45).
"""
import json
import os
import sys

REF = os.environ.get("REF", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))

# (test name, example deck, overrides from tests/<test>.sh, rtol used by the script)
CASES = [
    ("linear_wake.normalized.1Rank", "linear_wake",
     {"hipace.tile_size": "8", "diagnostic.field_data": "all rho"}, 1e-12),
    ("gaussian_linear_wake.normalized.1Rank", "linear_wake",
     {"hipace.tile_size": "8", "beam.profile": "gaussian", "beam.zmin": "-5.9", "beam.zmax": "5.9",
      "beam.radius": "10", "beam.position_mean": "0. 0. 0", "beam.position_std": "2 2 1.41",
      "geometry.prob_lo": "-10.   -10.   -6", "geometry.prob_hi": "10.    10.    6",
      "diagnostic.field_data": "all rho"}, 1e-12),
    ("beam_in_vacuum.normalized.1Rank", "beam_in_vacuum",
     {"hipace.tile_size": "8", "hipace.depos_order_xy": "0", "diagnostic.field_data": "all rho",
      "hipace.MG_tolerance_rel": "1e-5"}, 1e-12),
    ("blowout_wake.Serial", "blowout_wake", {"hipace.tile_size": "8"}, 2e-2),
    # max_step=1 on two ranks with dt = 0: the second step repeats the first one
    ("blowout_wake_explicit.2Rank", "blowout_wake", {"hipace.tile_size": "8", "max_step": "1"}, 1e-9),
    ("beam_in_vacuum.normalized.Serial", "beam_in_vacuum",
     {"hipace.tile_size": "8", "hipace.depos_order_xy": "0", "diagnostic.field_data": "all rho"}, 1e-9),
    # 21 steps (0..20) of a beam in vacuum in an external focusing field, beam pushed after
    # every slice; the checksums are those of the step-20 output
    ("beam_evolution.1Rank", "beam_in_vacuum",
     {"hipace.tile_size": "8", "amr.n_cell": "32 32 10", "max_step": "20", "geometry.prob_lo": "-2. -2. -2.",
      "geometry.prob_hi": "2.  2.  2.", "hipace.dt": "3.", "diagnostic.output_period": "20",
      "beam.density": "1.e-8", "beam.radius": "1.", "beam.ppc": "4 4 1",
      "beams.external_E(x,y,z,t)": ".5*x .5*y 0."}, 1e-12),
]


def parse_kv(text):
    out = {}
    for line in text.splitlines():
        line = line.split("#", 1)[0].strip()
        if line and "=" in line:
            k, v = line.split("=", 1)
            out[k.strip()] = v.strip()
    return out


def main():
    fixture = {}
    for test, example, overrides, rtol in CASES:
        deck = parse_kv(open(os.path.join(REF, "examples", example, "inputs_normalized")).read())
        golden = json.load(open(os.path.join(REF, "tests", "checksum", "benchmarks_json", test + ".json")))
        fixture[test] = {"deck": deck, "overrides": overrides, "rtol": rtol,
                         "fields": golden["lev=0"], "beam": golden.get("beam", {})}
    out = os.path.join(HERE, "reference_checksums.json")
    with open(out, "w") as f:
        json.dump(fixture, f, indent=1, sort_keys=True)
    print("wrote", out, file=sys.stderr)


if __name__ == "__main__":
    main()
