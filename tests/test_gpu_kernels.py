"""GPU parity of each per-slice operation against the oracle on identical seeded inputs.

Each test loads the same slice store and plasma state into the oracle (CPU FP64) and
into the HIP path (through the C ABI), runs one reference operation on both and
compares. The FP64 path with ordered deposition sums in the oracle's order (the reference
CPU build's tiled order), pairs DST rows as the oracle does and runs the multigrid and the
push in the same operation order: every result must agree to the last bit. The atomic
deposition sums in hardware order: within 1e-13 of the oracle norm-wise.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from helpers import PLASMA_REALS, random_plasma, random_store, rel_err, small_config  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _pair(cfg, n_particles=4000, seed=0, store_seed=1):
    from hipace2_amd.simulation import Simulation
    from oracle.oracle import OracleSim
    sim = Simulation(cfg)
    orc = OracleSim(cfg)
    st = random_store(sim.ncomp, sim.ldy, sim.ldx, seed=store_seed)
    sim.store.copy_(torch.as_tensor(st))
    orc.set_store(st)
    if n_particles:
        d = random_plasma(cfg, n_particles, seed=seed)
        orc.set_plasma(d)
        # resize the GPU plasma to the random state
        sim.np = n_particles
        for k in PLASMA_REALS:
            sim.plasma[k] = torch.as_tensor(d[k], dtype=torch.float64).to(sim.device)
        sim.plasma["valid"] = torch.as_tensor(d["valid"], dtype=torch.int32).to(sim.device)
        import hipace2_amd._abi as abi
        sim._pdesc = abi.Plasma(*[sim.plasma[k].data_ptr() for k in PLASMA_REALS + ["valid"]], n_particles)
    return sim, orc


# (deposition, bin_period, particles). ordered: about one particle per cell, and fifteen
# (lists of many particles merged per cell). atomic: index order; tile order with chunks
# wider than an LDS window (a mean-centred window plus direct atomics); dense tile order
# (whole chunks in LDS)
DEPOSIT_MODES = [("ordered", 0, 4000), ("ordered", 0, 60000), ("atomic", 0, 4000), ("atomic", 1, 4000),
                 ("atomic", 4, 60000)]


def _same(g, o, mode, name):
    """Bit-equal planes in ordered mode, within TOL norm-wise in atomic mode."""
    if mode == "ordered":
        assert np.array_equal(g, o), (name, rel_err(g, o))
    else:
        e = rel_err(g, o)
        assert e <= TOL, (name, e)


@pytest.mark.parametrize("mode", DEPOSIT_MODES)
@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_deposit_current(order, mode):
    cfg = small_config(order=order, dtype=2 if order > 0 else 1)
    cfg.deposition, cfg.bin_period, n = mode
    sim, orc = _pair(cfg, n_particles=n)
    comps = ["jx", "jy", -1, -1, "chi", "rhomjz"]
    for name in ("jx", "jy", "chi", "rhomjz"):
        sim.store[sim.layout[name]].zero_()
    st = orc.store()
    for name in ("jx", "jy", "chi", "rhomjz"):
        st[orc.layout[name]] = 0
    orc.set_store(st)
    nq_g = sim.deposit_current(comps)
    nq_o = orc.deposit_plasma(comps)
    assert nq_g == nq_o
    g, o = sim.store_host(), orc.store()
    for name in ("jx", "jy", "chi", "rhomjz"):
        _same(g[sim.layout[name]], o[orc.layout[name]], cfg.deposition, name)
    # QSA removals are applied to the same particles
    pg, po = sim.plasma_host(), orc.plasma()
    assert np.array_equal(pg["valid"], po["valid"])


@pytest.mark.parametrize("deposition", ["ordered", "atomic"])
def test_deposit_current_qsa_removal(deposition):
    cfg = small_config(order=2)
    cfg.deposition = deposition
    sim, orc = _pair(cfg, seed=3)
    d = orc.plasma()
    # force QSA violations: gamma/psi > 35 for psi -> small, and psi < 0
    d["psi"][:10] = 0.1
    d["psi"][10:20] = -0.5
    d["valid"][:20] = 1
    orc.set_plasma(d)
    sim.set_plasma(d)
    nq_g = sim.deposit_current(["jx", "jy", -1, -1, "chi", "rhomjz"])
    nq_o = orc.deposit_plasma(["jx", "jy", -1, -1, "chi", "rhomjz"])
    assert nq_g == nq_o and nq_g >= 20
    pg, po = sim.plasma_host(), orc.plasma()
    assert np.array_equal(pg["valid"], po["valid"])
    assert np.array_equal(pg["w"], po["w"])


@pytest.mark.parametrize("order,dtype", [(1, 2), (2, 2), (2, 1), (2, 0), (3, 2), (0, 1), (3, 1), (1, 0)])
@pytest.mark.parametrize("mode", DEPOSIT_MODES)
def test_explicit_deposition(order, dtype, mode):
    cfg = small_config(order=order, dtype=dtype)
    cfg.deposition, cfg.bin_period, n = mode
    sim, orc = _pair(cfg, n_particles=n)
    sim.explicit_deposition()
    orc.explicit_deposit()
    g, o = sim.store_host(), orc.store()
    for name in ("Sx", "Sy"):
        _same(g[sim.layout[name]], o[orc.layout[name]], cfg.deposition, name)


@pytest.mark.parametrize("order", [0, 1, 2, 3])
@pytest.mark.parametrize("deposition", ["ordered", "atomic"])
def test_deposit_beam(order, deposition):
    """A Gaussian beam slice whose core puts hundreds of particles into a cell (long lists),
    with jx, jy and jz."""
    from hipace2_amd.inputs import BeamSpec, make_beam
    cfg = small_config(order=order, dtype=2 if order > 0 else 1, nz=8)
    cfg.deposition = deposition
    spec = BeamSpec(injection_type="fixed_weight", profile="gaussian", density=1.0, position_mean=(0.1, -0.2, 0.0),
                    position_std=(0.25, 0.35, 1.0), num_particles=40000, u_mean=(0.3, -0.2, 500.0),
                    u_std=(1.0, 1.0, 5.0), seed=2)
    beam = make_beam(cfg, spec)
    sim, orc = _pair(cfg, n_particles=0)
    sim.set_beam(beam)
    orc.set_beam(beam)
    for islice in range(cfg.nz):
        sim.deposit_current_slice(islice, "jx_beam", "jy_beam", "jz_beam")
        orc.deposit_beam(islice, "jx_beam", "jy_beam", "jz_beam")
    g, o = sim.store_host(), orc.store()
    for name in ("jx_beam", "jy_beam", "jz_beam"):
        _same(g[sim.layout[name]], o[orc.layout[name]], deposition, name)


@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_advance_plasma_particles(order):
    cfg = small_config(order=order, dtype=2 if order > 0 else 1)
    for temp in (False, True):
        sim, orc = _pair(cfg, seed=5)
        sim.advance_plasma_particles(temp_slice=temp)
        orc.push_plasma(temp_slice=temp)
        pg, po = sim.plasma_host(), orc.plasma()
        for k in PLASMA_REALS:
            assert np.array_equal(pg[k], po[k]), (k, temp, rel_err(pg[k], po[k]))
        assert np.array_equal(pg["valid"], po["valid"])


@pytest.mark.parametrize("bc", ["Periodic", "Reflecting", "Absorbing"])
def test_push_boundary_conditions(bc):
    cfg = small_config(order=2)
    cfg.particle_bc = bc
    sim, orc = _pair(cfg, seed=7)
    d = orc.plasma()
    # kick a fraction of the particles across the boundary
    d["ux_half_step"][:200] = 40.0
    d["uy_half_step"][200:400] = -40.0
    orc.set_plasma(d)
    sim.set_plasma(d)
    sim.advance_plasma_particles()
    orc.push_plasma()
    pg, po = sim.plasma_host(), orc.plasma()
    for k in PLASMA_REALS + ["valid"]:
        assert np.array_equal(pg[k], po[k]), k


@pytest.mark.parametrize("n", [63, 64, 48, 255, 511, 1023])
def test_poisson_psi_ez_bz(n):
    """Three fields solved together, their rows paired across field boundaries (FFT path for
    n + 1 a power of two, direct sums otherwise): bit-equal."""
    cfg = small_config(n=n)
    sim, orc = _pair(cfg, n_particles=0)
    sim.solve_poisson_psi_exmby_eypbx_ez_bz()
    orc.solve_poisson_psi_ez_bz()
    g, o = sim.store_host(), orc.store()
    for name in ("Psi", "Ez", "Bz", "ExmBy", "EypBx"):
        assert np.array_equal(g[sim.layout[name]], o[orc.layout[name]]), (name, rel_err(g[sim.layout[name]],
                                                                                          o[orc.layout[name]]))


@pytest.mark.parametrize("n", [63, 64])
def test_poisson_bxby_predictor_corrector(n):
    cfg = small_config(n=n, explicit=False)
    sim, orc = _pair(cfg, n_particles=0)
    sim.solve_poisson_bxby()
    orc.solve_poisson_bxby()
    g, o = sim.store_host(), orc.store()
    for name in ("pciter_Bx", "pciter_By"):
        assert np.array_equal(g[sim.layout[name]], o[orc.layout[name]]), (name, rel_err(g[sim.layout[name]],
                                                                                          o[orc.layout[name]]))


# (tol_rel, maxiter): converged after odd and even numbers of V-cycles, converged before
# the first V-cycle (tolerance above the initial residual), stopped by maxiter (odd, even, 0)
MG_RUNS = [(1e-4, 200), (1e-7, 200), (1e-10, 200), (1.0, 200), (1e-300, 1), (1e-300, 2), (1e-300, 0)]
ORACLE_RC = {0: 0, 1: 3, 2: 4}


@pytest.mark.parametrize("n", [63, 64, 127, 32, 255, 511])
def test_mg_solve1_bit_exact(n):
    """Consecutive solves on one context (each starts from the previous solution) against
    the oracle: same return code, same V-cycle count, identical planes."""
    cfg = small_config(n=n)
    sim, orc = _pair(cfg, n_particles=0)
    st = orc.store()
    rng = np.random.default_rng(11)
    st[orc.layout["chi"]] = rng.uniform(0.0, 2.0, st[0].shape)
    orc.set_store(st)
    sim.store.copy_(torch.as_tensor(st))
    iters = set()
    for tol, maxiter in MG_RUNS:
        rc_g, it_g = sim.mg_solve1(tol_rel=tol, maxiter=maxiter)
        rc_o, it_o = orc.mg_solve1(tol_rel=tol, maxiter=maxiter)
        assert rc_g == ORACLE_RC[rc_o], (tol, maxiter, rc_g, rc_o)
        assert it_g == it_o, (tol, maxiter, it_g, it_o)
        iters.add(it_o)
        g, o = sim.store_host(), orc.store()
        for name in ("Bx", "By"):
            assert np.array_equal(g[sim.layout[name]], o[orc.layout[name]]), (
                tol, maxiter, name, rel_err(g[sim.layout[name]], o[orc.layout[name]]))
    assert 0 in iters and any(i % 2 == 1 for i in iters) and any(i % 2 == 0 and i > 0 for i in iters), iters


def test_mg_nonconvergence_reported():
    cfg = small_config(n=63)
    sim, orc = _pair(cfg, n_particles=0)
    rc, it = sim.mg_solve1(tol_rel=1e-300, maxiter=1)
    assert rc == 3 and it == 1
