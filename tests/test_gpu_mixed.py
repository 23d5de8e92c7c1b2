"""Mixed-precision path (precision="mixed"): FP32 gather, push increments and deposition
whose FP32 contributions are summed in FP64 LDS words, FP64 field solves and slice-to-slice
particle state.

Each operation is compared with the FP64 oracle on the same seeded inputs. The bounds
are error budgets, not parity: norm-wise relative error (max |a - b| / max |b|) of
  - deposition: per-particle FP32 quantities (a few roundings, ~5e-7 each) summed in
    FP64 window sums that add no error of their own at this level: 2e-6;
  - explicit deposition: fields rounded to FP32 and the Sx/Sy coefficients: 1e-5;
  - push: FP32 increments added to FP64 state: 1e-5 of the state's norm (the random
    fields drive a few particles to |u| ~ 90 and psi < 0, where six FP32 increments of
    O(10) each carry a few 1e-6 of the norm; cells and polynomial pieces are chosen in
    FP64, so particles sitting exactly on a piece boundary, as the clipped ones of the
    random state do, see the same stencil);
  - the whole explicit slice loop on a linear wake: Ez, Ex-By, jx and rho-jz (the
    explicit solver's jz carrier), Bx and By within 1e-4.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from helpers import PLASMA_REALS, rel_err, small_config  # noqa: E402
from test_gpu_kernels import DEPOSIT_MODES, _pair  # noqa: E402

pytestmark = pytest.mark.gpu


def _mixed(cfg):
    cfg.precision = "mixed"
    return cfg


@pytest.mark.parametrize("mode", DEPOSIT_MODES)
@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_mixed_deposit_current(order, mode):
    cfg = _mixed(small_config(order=order, dtype=2 if order > 0 else 1))
    cfg.deposition, cfg.bin_period, n = mode
    sim, orc = _pair(cfg, n_particles=n)
    comps = [k if sim.layout[k] >= 0 else -1 for k in ("jx", "jy", "jz", "rho", "chi", "rhomjz")]
    names = [k for k in comps if k != -1]
    for name in names:
        sim.store[sim.layout[name]].zero_()
    st = orc.store()
    for name in names:
        st[orc.layout[name]] = 0
    orc.set_store(st)
    assert sim.deposit_current(comps) == orc.deposit_plasma(comps)
    g, o = sim.store_host(), orc.store()
    for name in names:
        e = rel_err(g[sim.layout[name]], o[orc.layout[name]])
        assert e <= 2e-6, (name, e)
        assert e > 0.0 or order == 0, (name, "FP32 path expected to round")
    pg, po = sim.plasma_host(), orc.plasma()
    assert np.array_equal(pg["valid"], po["valid"])


@pytest.mark.parametrize("order", [1, 2])
def test_mixed_compensated_fp32_sums(order):
    """HPQ_SUM_KAHAN32: FP32 window words with an FP32 compensation word each. The FP32
    contributions are the same as with FP64 words; each addition's rounding error is recovered,
    so the planes stay within the FP32 contributions' budget (2e-6) and within a small factor of
    the FP64-word sums (the compensation word rounds in FP32 too: 2.8x on chi at order 1)."""
    errs = {}
    for sums in ("fp64", "kahan32"):
        cfg = _mixed(small_config(order=order, dtype=2))
        cfg.deposition, cfg.bin_period = "atomic", 1
        cfg.mixed_sums = sums
        sim, orc = _pair(cfg, n_particles=60000)
        comps = [k if sim.layout[k] >= 0 else -1 for k in ("jx", "jy", "jz", "rho", "chi", "rhomjz")]
        names = [k for k in comps if k != -1]
        for name in names:
            sim.store[sim.layout[name]].zero_()
        st = orc.store()
        for name in names:
            st[orc.layout[name]] = 0
        orc.set_store(st)
        assert sim.deposit_current(comps) == orc.deposit_plasma(comps)
        sim.explicit_deposition()
        orc.explicit_deposit()
        g, o = sim.store_host(), orc.store()
        errs[sums] = {name: rel_err(g[sim.layout[name]], o[orc.layout[name]]) for name in names + ["Sx", "Sy"]}
    for name, e in errs["kahan32"].items():
        assert e <= 2e-6, (name, e)
        assert e <= 10.0 * errs["fp64"][name] + 1e-9, (name, e, errs["fp64"][name])


@pytest.mark.parametrize("order,dtype", [(1, 2), (2, 2), (2, 1), (2, 0), (3, 2), (0, 1)])
@pytest.mark.parametrize("mode", DEPOSIT_MODES)
def test_mixed_explicit_deposition(order, dtype, mode):
    cfg = _mixed(small_config(order=order, dtype=dtype))
    cfg.deposition, cfg.bin_period, n = mode
    sim, orc = _pair(cfg, n_particles=n)
    sim.explicit_deposition()
    orc.explicit_deposit()
    g, o = sim.store_host(), orc.store()
    for name in ("Sx", "Sy"):
        e = rel_err(g[sim.layout[name]], o[orc.layout[name]])
        assert e <= 1e-5, (name, e)


@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_mixed_push(order):
    cfg = _mixed(small_config(order=order, dtype=2 if order > 0 else 1))
    for temp in (False, True):
        sim, orc = _pair(cfg, seed=5)
        sim.advance_plasma_particles(temp_slice=temp)
        orc.push_plasma(temp_slice=temp)
        pg, po = sim.plasma_host(), orc.plasma()
        for k in PLASMA_REALS:
            e = rel_err(pg[k], po[k])
            if e > 1e-5:
                w = int(np.argmax(np.abs(pg[k] - po[k])))
                print("worst particle", k, w, {q: (float(pg[q][w]), float(po[q][w])) for q in PLASMA_REALS + ["valid"]})
            assert e <= 1e-5, (k, temp, e)
        assert np.array_equal(pg["valid"], po["valid"])


@pytest.mark.parametrize("n,order,peak,guess", [(63, 2, 1.0, "previous"), (64, 1, 0.5, "previous"),
                                               (63, 2, 1.0, "dst")])
def test_mixed_slices_vs_fp64_oracle(n, order, peak, guess):
    """The mixed path against the FP64 oracle, every slice within 1e-4; also with the
    multigrid started from the DST-corrected guess (the benchmark's mixed path). That guess
    changes where the multigrid stops inside its tolerance (1e-4 of the residual): Ez stays
    within 1e-4, the other fields within 3e-4 (FP32 arithmetic alone stays near 1e-6 here)."""
    from hipace2_amd.inputs import BeamSpec, make_beam
    from hipace2_amd.simulation import Simulation
    from oracle.oracle import OracleSim
    cfg = _mixed(small_config(n=n, nz=48, order=order, explicit=True))
    cfg.mg_guess = guess
    spec = BeamSpec(injection_type="fixed_weight", profile="gaussian", density=peak,
                    position_mean=(0.0, 0.0, 0.0), position_std=(0.3, 0.3, 1.41),
                    num_particles=20000, u_mean=(0.0, 0.0, 1000.0), seed=0)
    beam = make_beam(cfg, spec)
    sim = Simulation(cfg, beam)
    orc = OracleSim(cfg, beam)
    names = [k for k in ("Ez", "ExmBy", "jx", "rhomjz", "Bx", "By") if sim.layout[k] >= 0]
    cg = {k: sim.capture(k) for k in names}
    co = {k: orc.capture(k) for k in names}
    sim.run_step()
    orc.run_step()
    torch.cuda.synchronize()
    errs = {k: rel_err(cg[k].cpu().numpy(), co[k]) for k in names}
    print("mixed vs FP64 oracle:", errs)
    for k, e in errs.items():
        assert e <= (1e-4 if guess == "previous" or k == "Ez" else 3e-4), (k, e)


# Split-BF16 sine-matrix DST (poisson_solver="bf16x3"): operands carry ~16 significant bits
# (hi + lo), products are accumulated in FP32 over K <= 512 terms; four products per
# solve. Bound: 1e-4 of the field's norm against the FP64 FFT solve of the oracle (the
# right-hand sides of Ez and Bz here are differences of random currents, white noise whose
# solution is small next to the operands, the hardest case for a relative bound).
BF16X3_TOL = 1e-4


@pytest.mark.parametrize("n", [63, 64, 48, 255, 511])
def test_poisson_bf16x3(n):
    cfg = small_config(n=n)
    cfg.poisson_solver = "bf16x3"
    sim, orc = _pair(cfg, n_particles=0)
    sim.solve_poisson_psi_exmby_eypbx_ez_bz()
    orc.solve_poisson_psi_ez_bz()
    g, o = sim.store_host(), orc.store()
    # norm-wise over the written region (valid box; plus one guard ring for ExmBy, EypBx),
    # not over the untouched random guard values
    box = {"Psi": 2, "Ez": 2, "Bz": 2, "ExmBy": 1, "EypBx": 1}
    errs = {name: rel_err(g[sim.layout[name]][r:-r, r:-r], o[orc.layout[name]][r:-r, r:-r])
            for name, r in box.items()}
    print("bf16x3 vs FP64 FFT:", errs)
    for name, e in errs.items():
        assert e <= BF16X3_TOL, (name, e)


def test_poisson_bf16x3_rectangular_and_pc():
    cfg = small_config(n=63, explicit=False)
    cfg.ny = 101
    cfg.poisson_solver = "bf16x3"
    sim, orc = _pair(cfg, n_particles=0)
    sim.solve_poisson_bxby()
    orc.solve_poisson_bxby()
    g, o = sim.store_host(), orc.store()
    for name in ("pciter_Bx", "pciter_By"):
        e = rel_err(g[sim.layout[name]][2:-2, 2:-2], o[orc.layout[name]][2:-2, 2:-2])
        assert e <= BF16X3_TOL, (name, e)


@pytest.mark.parametrize("nx,ny", [(63, 63), (511, 511), (48, 48), (63, 101)])
def test_poisson_bf16x3_matches_restatement(nx, ny):
    """hpq_poisson_dirichlet on the bf16x3 path against the numpy restatement of the same
    splitting: they differ only in the order of the FP32 accumulation inside each product."""
    from hipace2_amd.simulation import Simulation
    from oracle import oracle as orc
    cfg = small_config(n=nx)
    cfg.ny = ny
    cfg.poisson_solver = "bf16x3"
    sim = Simulation(cfg)
    S = np.random.default_rng(9).normal(size=(ny, nx))
    lhs = torch.zeros((sim.ldy, sim.ldx), dtype=torch.float64, device=sim.device)
    sim.poisson_dirichlet(torch.as_tensor(S).to(sim.device).contiguous(), lhs)
    got = lhs.cpu().numpy()
    want = orc.poisson_bf16x3(cfg, S)
    e = rel_err(got[2:-2, 2:-2], want)
    print("bf16x3 GPU vs restatement:", e)
    assert e <= 1e-5, e
    assert np.all(got[:2] == 0) and np.all(got[-2:] == 0) and np.all(got[:, :2] == 0) and np.all(got[:, -2:] == 0)
