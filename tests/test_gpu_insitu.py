"""In-situ reduced diagnostics on the device (hpq_set_insitu) against the oracle's NumPy
restatement (oracle/insitu.py) applied to the same states: the field planes captured at the
same point of each slice, the plasma as it stands at the start of each slice, the beam
particles of each slice before their push. The device sums in another order than numpy,
so values agree to rounding: 1e-12 of each row's scale. Particle counts are exact."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from helpers import small_config  # noqa: E402

pytestmark = pytest.mark.gpu

FIELD_PLANES = ("ExmBy", "EypBx", "Ez", "Bx", "By", "Bz", "jz_beam")


def _beam(cfg, n=20000, seed=5, uz=20.0):
    from hipace2_amd.inputs import BeamSpec, make_beam
    spec = BeamSpec(injection_type="fixed_weight", profile="gaussian", density=0.5,
                    position_mean=(0.1, -0.05, 0.0), position_std=(0.4, 0.3, 1.41), num_particles=n,
                    u_mean=(0.2, 0.0, uz), u_std=(0.5, 0.4, 1.0), seed=seed)
    return make_beam(cfg, spec)


def _close_rows(got, want, what):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    for k in range(want.shape[0]):
        scale = max(1e-300, float(np.max(np.abs(want[k]))))
        err = float(np.max(np.abs(got[k] - want[k]))) / scale
        assert err <= 1e-12, (what, k, err)


def _oracle_beam(beam, off, nz, radius=math.inf):
    from oracle import insitu as orc
    per = np.zeros((22, nz))
    sums = np.zeros(22)
    npart = np.zeros(nz, np.int64)
    for s in range(nz):
        a, b = int(off[s]), int(off[s + 1])
        sl = slice(a, b)
        p, sm, n = orc.beam(beam["x"][sl], beam["y"][sl], beam["z"][sl], beam["ux"][sl], beam["uy"][sl],
                            beam["uz"][sl], beam["w"][sl], beam["valid"][sl], radius)
        per[:, s], npart[s] = p, n
        sums += sm
    return per, sums, npart


def _host_beam(b):
    return {k: np.asarray(getattr(b, k)) for k in ("x", "y", "z", "ux", "uy", "uz", "w", "valid")}


@pytest.mark.parametrize("radius", [math.inf, 2.0])
def test_insitu_slice_by_slice_matches_oracle(radius):
    from hipace2_amd.simulation import Simulation
    from oracle import insitu as orc
    cfg = small_config(n=63, nz=16, order=2, explicit=True)
    beam = _beam(cfg)
    sim = Simulation(cfg, beam)
    sim.enable_insitu(plasma_radius=radius, beam_radius=radius)
    caps = {c: sim.capture(c) for c in FIELD_PLANES}
    sim.begin_step()
    nz = cfg.nz
    pl_per = np.zeros((14, nz))
    pl_sum = np.zeros(14)
    pl_np = np.zeros(nz, np.int64)
    for isl in range(nz - 1, -1, -1):
        p = sim.plasma_host()
        per, sm, n = orc.plasma(p["x"], p["y"], p["w"], p["ux"], p["uy"], p["psi"], p["valid"], radius)
        pl_per[:, isl], pl_np[isl] = per, n
        pl_sum += sm
        sim.solve_slice(isl)
    a = sim.insitu_arrays()
    # fields
    planes = {c: caps[c].cpu().numpy() for c in FIELD_PLANES}
    f_per = np.zeros((10, nz))
    for s in range(nz):
        f_per[:, s], _ = orc.fields(*[planes[c][s] for c in FIELD_PLANES], cfg.dx, cfg.dy, cfg.dz)
    _close_rows(a["field_rdata"], f_per, "fields")
    np.testing.assert_allclose(a["field_sum"], f_per.sum(axis=1), rtol=1e-12,
                               atol=1e-12 * float(np.abs(f_per).sum()))
    assert np.any(f_per[8] != 0) and np.any(f_per[9] != 0)
    # plasma
    assert np.array_equal(a["plasma_idata"], pl_np) and int(a["plasma_isum"][0]) == int(pl_np.sum())
    _close_rows(a["plasma_rdata"], pl_per, "plasma")
    np.testing.assert_allclose(a["plasma_sum"], pl_sum, rtol=1e-12, atol=1e-12 * float(np.abs(pl_sum).max()))
    # beam: the particles of each slice, unmoved by the static-beam loop
    b_per, b_sum, b_np = _oracle_beam(_host_beam(beam), beam.off, nz, radius)
    assert np.array_equal(a["beam_idata"], b_np) and int(a["beam_isum"][0]) == int(b_np.sum())
    assert b_np.sum() > 1000
    _close_rows(a["beam_rdata"], b_per, "beam")
    np.testing.assert_allclose(a["beam_sum"], b_sum, rtol=1e-12, atol=1e-12 * float(np.abs(b_sum).max()))


def test_insitu_moving_beam_and_files(tmp_path):
    """hpq_run_step_with_beam: each slice's beam moments are taken before that slice's push,
    so they are the moments of the step's initial beam; two steps write two records."""
    from hipace2_amd import insitu as io
    from hipace2_amd.simulation import Simulation
    from oracle import insitu as orc
    cfg = small_config(n=63, nz=24, order=2, explicit=True)
    beam = _beam(cfg, seed=9)
    sim = Simulation(cfg, beam)
    sim.enable_insitu()
    caps = {c: sim.capture(c) for c in FIELD_PLANES}
    written = []
    for step in range(2):
        if step > 0:
            sim.sort_beam()
        torch.cuda.synchronize()
        b0 = {k: v[: beam.n].cpu().numpy() for k, v in sim.beam.items() if k != "nsub"}
        off = np.asarray(sim.beam_host.off)
        sim.run_step_with_beam(4.0, 10)
        a = sim.insitu_arrays()
        b_per, b_sum, b_np = _oracle_beam(b0, off, cfg.nz)
        assert np.array_equal(a["beam_idata"], b_np), step
        _close_rows(a["beam_rdata"], b_per, ("beam", step))
        planes = {c: caps[c].cpu().numpy() for c in FIELD_PLANES}
        f_per = np.stack([orc.fields(*[planes[c][s] for c in FIELD_PLANES], cfg.dx, cfg.dy, cfg.dz)[0]
                          for s in range(cfg.nz)], axis=1)
        _close_rows(a["field_rdata"], f_per, ("fields", step))
        written.append(a)
        sim.write_insitu(step, 4.0 * step, field_prefix=str(tmp_path / "f"), plasma_prefix=str(tmp_path / "p"),
                         beam_prefix=str(tmp_path / "b"))
        after = sim.insitu_arrays()
        assert all(not np.any(v) for v in after.values())
    f = io.read_file(str(tmp_path / "f" / "reduced_fields.*.txt"))
    p = io.read_file(str(tmp_path / "p" / "reduced_plasma.*.txt"))
    b = io.read_file(str(tmp_path / "b" / "reduced_beam.*.txt"))
    assert f.shape == p.shape == b.shape == (2,)
    for step, a in enumerate(written):
        assert f[step]["step"] == step and f[step]["time"] == 4.0 * step and f[step]["n_slices"] == cfg.nz
        for i, n in enumerate(io.FIELD_NAMES):
            np.testing.assert_array_equal(f[step][n], a["field_rdata"][i])
            assert f[step]["integrated"][n] == a["field_sum"][i]
        for i, n in enumerate(io.BEAM_NAMES[1:], start=1):
            np.testing.assert_array_equal(b[step][n], a["beam_rdata"][i])
            assert b[step]["average"][n] == a["beam_sum"][i] / a["beam_sum"][0]
        np.testing.assert_array_equal(b[step]["Np"], a["beam_idata"])
        np.testing.assert_array_equal(p[step]["Np"], a["plasma_idata"])
        assert p[step]["total"]["[(ga-1)*(1-vz)]"] == a["plasma_sum"][13]
        assert b[step]["charge"] == beam.charge and p[step]["normalized_density_factor"] == cfg.dx * cfg.dy * cfg.dz


def test_insitu_predictor_corrector_and_errors():
    """Field diagnostics need the explicit solver's jz_beam plane (Fields.cpp:1313-1314);
    the particle diagnostics do not."""
    from hipace2_amd import _abi
    from hipace2_amd.simulation import Simulation
    cfg = small_config(n=63, nz=8, order=2, explicit=False)
    beam = _beam(cfg, n=5000)
    sim = Simulation(cfg, beam)
    with pytest.raises(_abi.HpqError) as e:
        sim.enable_insitu(fields=True)
    assert _abi.ERRORS[e.value.code] == "HPQ_E_UNSUPPORTED"
    assert "explicit solver" in str(e.value)
    with pytest.raises(_abi.HpqError):
        sim.enable_insitu(fields=False, plasma_radius=-1.0)
    sim.enable_insitu(fields=False)
    sim.run_step()
    a = sim.insitu_arrays()
    b_per, _, b_np = _oracle_beam(_host_beam(beam), beam.off, cfg.nz)
    assert np.array_equal(a["beam_idata"], b_np)
    _close_rows(a["beam_rdata"], b_per, "beam")
    assert a["plasma_idata"].min() > 0
    sim.disable_insitu()
    sim.run_step()
    assert sim.insitu_arrays() == {}


def test_insitu_not_repeated_by_adjoint_recompute():
    """The adjoint's backward pass recomputes slices; the diagnostics of the forward pass
    are neither overwritten nor added twice."""
    from hipace2_amd.adjoint import slice_loop_ez
    from hipace2_amd.simulation import Simulation
    cfg = small_config(n=63, nz=20, order=2, explicit=True)
    cfg.do_beam_jx_jy = False
    beam = _beam(cfg, n=10000, uz=1000.0)
    sim = Simulation(cfg, beam)
    sim.enable_insitu()
    sim.run_step()
    ref = sim.insitu_arrays()
    sim.reset_insitu()
    w = sim.beam["w"][: beam.n].clone().requires_grad_(True)
    ez = slice_loop_ez(w, sim, checkpoint_every=6)
    ez.sum().backward()
    got = sim.insitu_arrays()
    for k in ref:
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-9, atol=1e-12 * max(1.0, float(np.abs(ref[k]).max())),
                                   err_msg=k)
