"""In-situ diagnostics on the CPU: the file format (header text and binary records of
src/utils/InsituUtil.H:23-87) and the oracle's restatement of the reductions against closed
forms and brute force on a few particles."""
import math
import os

import numpy as np
import pytest

from hipace2_amd import insitu as io
from oracle import insitu as orc


def test_header_layout_known_answer():
    nodes = [io._leaf("time", 1.5, "<f8"), io._leaf("step", 3, "<i4"), io._leaf("arr", [1.0, 2.0, 3.0], "<f8"),
             ("integrated", [io._leaf("b", 4.0, "<f8")])]
    expected = (
        '{\n'
        '    "names": [\n'
        '        "time",\n'
        '        "step",\n'
        '        "arr",\n'
        '        "integrated"\n'
        '    ],\n'
        '    "formats": [\n'
        '        "<f8",\n'
        '        "<i4",\n'
        '        "(3,)<f8",\n'
        '        {\n'
        '            "names": [\n'
        '                "b"\n'
        '            ],\n'
        '            "formats": [\n'
        '                "<f8"\n'
        '            ]\n'
        '        }\n'
        '    ]\n'
        '}')
    assert io.header(nodes) == expected
    raw = io.data(nodes)
    assert len(raw) == 8 + 4 + 24 + 8
    assert np.frombuffer(raw[12:36], "<f8").tolist() == [1.0, 2.0, 3.0]
    assert np.frombuffer(raw[:8], "<f8")[0] == 1.5 and np.frombuffer(raw[8:12], "<i4")[0] == 3


def test_records_append_and_read_back(tmp_path):
    rng = np.random.default_rng(0)
    nz = 5
    path = io.file_name(str(tmp_path), "fields", 7)
    assert os.path.basename(path) == "reduced_fields.0007.txt"
    recs = []
    for step, t in ((2, 0.4), (0, 0.0), (1, 0.2)):
        rd, s = rng.normal(size=(10, nz)), rng.normal(size=10)
        io.append_record(path, io.field_record(t, step, -8.0, 4.0, rd, s))
        recs.append((step, t, rd, s))
    allr = io.read_file(str(tmp_path / "reduced_fields.*.txt"))
    assert allr.shape == (3,) and list(allr["step"]) == [0, 1, 2]
    for step, t, rd, s in recs:
        r = allr[allr["step"] == step][0]
        assert r["time"] == t and r["n_slices"] == nz and r["is_normalized_units"] == 1
        assert r["z_lo"] == -8.0 and r["z_hi"] == 4.0
        for i, n in enumerate(io.FIELD_NAMES):
            np.testing.assert_array_equal(r[n], rd[i])
            assert r["integrated"][n] == s[i]


@pytest.mark.parametrize("kind,nz", [("plasma", 4), ("beam", 4), ("beam", 1)])
def test_species_record_fields(tmp_path, kind, nz):
    names = io.PLASMA_NAMES if kind == "plasma" else io.BEAM_NAMES
    rng = np.random.default_rng(1)
    rd, idata = rng.normal(size=(len(names), nz)), rng.integers(0, 100, nz).astype(np.int32)
    sums = rng.uniform(1, 2, size=len(names))
    path = io.file_name(str(tmp_path), kind, 0)
    io.append_record(path, io.species_record(kind, 0.5, 4, -1.0, 1.0, -8.0, 4.0, 0.125, rd, idata, sums, 77))
    r = io.read_file(path)[0]
    assert r["charge"] == -1.0 and r["mass"] == 1.0 and r["normalized_density_factor"] == 0.125
    np.testing.assert_array_equal(np.atleast_1d(r["sum(w)"]), rd[0])
    np.testing.assert_array_equal(np.atleast_1d(r["Np"]), idata)
    assert r["total"]["sum(w)"] == sums[0] and r["total"]["Np"] == 77
    n_avg = 13 if kind == "plasma" else len(names)
    for i in range(1, n_avg):
        np.testing.assert_array_equal(np.atleast_1d(r[names[i]]), rd[i])
        assert r["average"][names[i]] == sums[i] / sums[0]
    if kind == "plasma":
        assert r["total"]["[(ga-1)*(1-vz)]"] == sums[13]
        assert "[(ga-1)*(1-vz)]" not in r["average"].dtype.names
    # one slice: plain scalars, no sub-array, as the reference's "size > 1" rule gives
    if nz == 1:
        assert '"(1,)' not in io.header(io.species_record(kind, 0, 0, 1, 1, 0, 1, 1, rd, idata, sums, 1))


# ---- the oracle's reductions --------------------------------------------------------------

def test_oracle_fields_constant_planes():
    ny, nx, dx, dy, dz = 3, 4, 0.5, 0.25, 0.1
    one = np.ones((ny, nx))
    a, b, ez, bx, by, bz, jz = 0.3, -1.1, 2.0, 0.7, -0.2, 0.05, -3.0
    per, s = orc.fields(a * one, b * one, ez * one, bx * one, by * one, bz * one, jz * one, dx, dy, dz)
    vol = nx * ny * dx * dy * dz
    want = [(a + by) ** 2, (b - bx) ** 2, ez ** 2, bx ** 2, by ** 2, bz ** 2, a ** 2, b ** 2, jz, ez * jz]
    np.testing.assert_allclose(per, np.array(want) * vol, rtol=1e-14)
    np.testing.assert_array_equal(per, s)


def test_oracle_plasma_closed_forms():
    rng = np.random.default_rng(3)
    n = 200
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    ux, uy = rng.normal(0, 0.3, n), rng.normal(0, 0.3, n)
    psi = rng.uniform(0.5, 1.5, n)
    w = rng.uniform(0.5, 2, n)
    valid = np.ones(n, np.int32)
    valid[::7] = 0
    per, s, npart = orc.plasma(x, y, w, ux, uy, psi, valid)
    m = valid != 0
    assert npart == m.sum()
    # the quasi-static gamma is the Lorentz factor of (ux, uy, uz) with uz = gamma - psi
    ga = (1 + ux ** 2 + uy ** 2 + psi ** 2) / (2 * psi)
    uz = ga - psi
    np.testing.assert_allclose(ga ** 2, 1 + ux ** 2 + uy ** 2 + uz ** 2, rtol=1e-13)
    wq = (w * ga / psi)[m]
    assert s[0] == pytest.approx(wq.sum(), rel=1e-14)
    assert per[1] == pytest.approx((wq * x[m]).sum() / wq.sum(), rel=1e-13)
    # [ga^2] - [uz^2] - [ux^2] - [uy^2] = 1 for the quasi-static weighted averages
    assert per[12] - per[10] - per[6] - per[8] == pytest.approx(1.0, rel=1e-12)
    assert s[13] == pytest.approx((w * (ga - 1))[m].sum(), rel=1e-13)
    assert per[13] == s[13] and per[0] == s[0]
    # particles at rest: gamma 1, no energy, uz 0
    per0, s0, _ = orc.plasma(x, y, w, 0 * x, 0 * x, np.ones(n), np.ones(n))
    assert per0[11] == pytest.approx(1.0, rel=1e-15) and s0[13] == 0.0 and per0[9] == 0.0
    assert s0[0] == pytest.approx(w.sum())


def test_oracle_plasma_radius_and_empty():
    x = np.array([0.0, 3.0, 0.5])
    y = np.array([0.0, 4.0, 0.0])
    one = np.ones(3)
    per, s, npart = orc.plasma(x, y, one, 0 * one, 0 * one, one, one, radius=5.0)
    assert npart == 3  # r = 5 exactly is kept (only r^2 > radius^2 is cut)
    per, s, npart = orc.plasma(x, y, one, 0 * one, 0 * one, one, one, radius=4.999)
    assert npart == 2 and per[1] == pytest.approx(0.25)
    per, s, npart = orc.plasma(x, y, one, 0 * one, 0 * one, one, 0 * one)
    assert npart == 0 and np.all(per == 0) and np.all(s == 0)


def test_oracle_beam_two_particles():
    x, y, z = np.array([1.0, -1.0]), np.array([0.5, 0.5]), np.array([2.0, 4.0])
    ux, uy, uz = np.array([0.1, -0.1]), np.array([0.0, 0.2]), np.array([10.0, 0.0])
    w = np.array([1.0, 3.0])
    per, s, npart = orc.beam(x, y, z, ux, uy, uz, w, np.ones(2))
    assert npart == 2 and per[0] == 4.0
    assert per[1] == pytest.approx((1.0 - 3.0) / 4)                 # [x]
    assert per[5] == pytest.approx((2.0 + 12.0) / 4)                # [z]
    assert per[13] == pytest.approx((1 * 1.0 * 0.1 + 3 * -1.0 * -0.1) / 4)  # [x*ux]
    assert per[18] == pytest.approx(0.1 / 10.0 / 4)                 # [ux/uz]: uz = 0 contributes 0
    g = np.sqrt(1 + ux ** 2 + uy ** 2 + uz ** 2)
    assert per[20] == pytest.approx((g * w).sum() / 4) and per[21] == pytest.approx((g * g * w).sum() / 4)
    # a linear emittance-free beam: x proportional to ux gives [x^2][ux^2] = [x*ux]^2
    rng = np.random.default_rng(5)
    xx = rng.normal(size=50)
    p2, _, _ = orc.beam(xx, 0 * xx, 0 * xx, 3 * xx, 0 * xx, 1 + 0 * xx, np.ones(50), np.ones(50))
    emit2 = (p2[2] - p2[1] ** 2) * (p2[8] - p2[7] ** 2) - (p2[13] - p2[1] * p2[7]) ** 2
    assert abs(emit2) <= 1e-12 * p2[2] * p2[8]
    assert math.isclose(s[0], 4.0)
