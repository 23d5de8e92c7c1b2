"""Closed-form and brute-force pins of the oracle's components, each chosen so that a
plausible mistake (dropped term, wrong sign or index, transposed operand) fails it."""
import math

import numpy as np
import pytest
import scipy.fft
import scipy.sparse
import scipy.sparse.linalg

from helpers import PLASMA_REALS, rel_err, small_config
from oracle import oracle as orc


# ----------------------------------------------------------------------------- DST / Poisson

@pytest.mark.parametrize("n", [1, 2, 3, 7, 31, 32, 63, 100, 255])
def test_dst1_is_fftw_rodft00(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n)
    want = scipy.fft.dst(x, type=1)  # y_k = 2 sum x_j sin(pi (j+1)(k+1)/(n+1)) = FFTW RODFT00
    assert rel_err(orc.dst1(x), want) < 1e-14


def _lap5(F, dx, dy):
    """5-point Laplacian of the valid region with zero values one cell outside."""
    P = np.pad(F, 1)
    return ((P[1:-1, 2:] - 2 * P[1:-1, 1:-1] + P[1:-1, :-2]) / dx ** 2 +
            (P[2:, 1:-1] - 2 * P[1:-1, 1:-1] + P[:-2, 1:-1]) / dy ** 2)


@pytest.mark.parametrize("n", [15, 32, 63])
def test_poisson_inverts_the_discrete_laplacian(n):
    cfg = small_config(n=n)
    cfg.ny = n + 4 if n % 2 else n  # non-square too
    rng = np.random.default_rng(3)
    S = rng.normal(size=(cfg.ny, cfg.nx))
    lhs = orc.poisson(cfg, S)
    F = lhs[2:-2, 2:-2]
    assert rel_err(_lap5(F, cfg.dx, cfg.dy), S) < 1e-12
    # guards untouched
    assert np.all(lhs[:2] == 0) and np.all(lhs[:, :2] == 0)


def test_poisson_eigenmode():
    cfg = small_config(n=31)
    j, i = np.meshgrid(np.arange(cfg.ny), np.arange(cfg.nx), indexing="ij")
    kx, ky = 3, 5
    mode = np.sin(np.pi * kx * (i + 1) / (cfg.nx + 1)) * np.sin(np.pi * ky * (j + 1) / (cfg.ny + 1))
    lam = -4 * (np.sin(np.pi * kx / (2 * (cfg.nx + 1))) ** 2 / cfg.dx ** 2 +
                np.sin(np.pi * ky / (2 * (cfg.ny + 1))) ** 2 / cfg.dy ** 2)
    lhs = orc.poisson(cfg, mode)
    assert rel_err(lhs[2:-2, 2:-2], mode / lam) < 1e-13


@pytest.mark.parametrize("nx,ny", [(63, 63), (64, 64), (48, 48), (63, 101)])
def test_poisson_bf16x3_error_budget(nx, ny):
    """The split-BF16 restatement against the FP64 solve: within 5e-5 of the norm, and
    not exact (the BF16 rounding is really there)."""
    cfg = small_config(n=nx)
    cfg.ny = ny
    S = np.random.default_rng(5).normal(size=(ny, nx))
    want = orc.poisson(cfg, S)[2:-2, 2:-2]
    got = orc.poisson_bf16x3(cfg, S)
    e = rel_err(got, want)
    assert 1e-7 < e < 5e-5, e
    # the residual of the 5-point Laplacian stays at the same level
    assert rel_err(_lap5(got, cfg.dx, cfg.dy), S) < 1e-3


def test_bf16_round_to_nearest_even():
    # 1 + 2^-8 is halfway between two BF16 values: ties go to the even mantissa (1.0);
    # 1 + 3 * 2^-8 goes up to 1 + 2^-6
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.0e-3], dtype=np.float32)
    r = orc.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 2 ** -8 * 3.0e-3  # half a unit in the 8th significant bit
    hi, lo = orc.split_bf16(np.float32(math.pi))
    assert abs(float(hi) + float(lo) - float(np.float32(math.pi))) < 2 ** -16 * math.pi


# ----------------------------------------------------------------------------- shape factors

def bspline(o, u):
    """Centred cardinal B-spline of order o (support width o+1), written independently."""
    a = abs(u)
    if o == 0:
        return 1.0 if a < 0.5 else (0.5 if a == 0.5 else 0.0)
    if o == 1:
        return max(0.0, 1.0 - a)
    if o == 2:
        if a < 0.5:
            return 0.75 - a * a
        if a < 1.5:
            return 0.5 * (1.5 - a) ** 2
        return 0.0
    if o == 3:
        if a < 1:
            return 2 / 3 - a * a + a ** 3 / 2
        if a < 2:
            return (2 - a) ** 3 / 6
        return 0.0
    if o == 4:
        if a < 0.5:
            return 115 / 192 - 5 * a * a / 8 + a ** 4 / 4
        if a < 1.5:
            return (55 + 20 * a - 120 * a * a + 80 * a ** 3 - 16 * a ** 4) / 96
        if a < 2.5:
            return (5 - 2 * a) ** 4 / 384
        return 0.0
    raise ValueError(o)


def dbspline(o, u, h=1e-6):
    return (bspline(o, u + h) - bspline(o, u - h)) / (2 * h)


XS = [0.0, 0.25, 0.49, 0.5, 0.51, 0.75, 3.1, -2.37, 10.999]


@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_shape_is_bspline_and_partition_of_unity(order):
    for xm in XS:
        tot, mom = 0.0, 0.0
        for ix in range(order + 1):
            s, cell = orc.shape(order, xm, ix)
            u = xm - cell
            if abs(abs(u) - 0.5) > 1e-12:
                assert abs(s - bspline(order, u)) < 1e-14, (order, xm, ix)
            elif order == 0:
                assert s == 1.0  # nearest-grid-point: a tie goes to the upper cell
            tot += s
            mom += s * (cell - xm)
        assert abs(tot - 1) < 1e-14
        if order >= 1:
            assert abs(mom) < 1e-13


@pytest.mark.parametrize("dtype", [0, 1, 2])
@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_derivative_shape_factors(dtype, order):
    if dtype == 0 and order == 0:
        pytest.skip("analytic derivative of order 0 vanishes (the reference forbids it)")
    width = order + 1 + dtype
    for xm in [0.1, 0.3, 0.45, 0.55, 0.8, -3.3, 7.6]:
        tot_s = tot_d = mom_d = 0.0
        for ix in range(width):
            s, ds, cell = orc.dshape(dtype, order, xm, ix)
            u = xm - cell
            assert abs(s - bspline(order, u)) < 1e-14, (dtype, order, xm, ix)
            if dtype == 0:
                want = -dbspline(order, u)
            elif dtype == 1:
                want = -dbspline(order + 1, u)
            else:
                want = -(bspline(order, u + 1) - bspline(order, u - 1)) / 2
            assert abs(ds - want) < 2e-9, (dtype, order, xm, ix, ds, want)
            tot_s += s
            tot_d += ds
            mom_d += ds * (cell - xm)
        assert abs(tot_s - 1) < 1e-14
        assert abs(tot_d) < 1e-14
        # ds is minus the derivative: sum_k ds_k (k - x) = -1, i.e. the derivative stencil
        # differentiates linear functions exactly
        assert abs(mom_d + 1) < 1e-13


# ----------------------------------------------------------------------------- multigrid

def _operator(cfg, chi):
    """Sparse matrix of -chi*u + Lap(u) on the MG unknowns (node: slice cells, Dirichlet 0
    one cell out; cell-centred: 4/3 ghost stencil), built independently of the oracle."""
    nx, ny = cfg.nx, cfg.ny
    fx, fy = 1 / cfg.dx ** 2, 1 / cfg.dy ** 2
    cc = nx % 2 == 0
    N = nx * ny
    A = scipy.sparse.lil_matrix((N, N))
    idx = lambda i, j: j * nx + i
    for j in range(ny):
        for i in range(nx):
            r = idx(i, j)
            A[r, r] += -chi[j, i] - 2 * fx - 2 * fy
            for (di, dj, f, n, k) in ((-1, 0, fx, nx, i), (1, 0, fx, nx, i), (0, -1, fy, ny, j), (0, 1, fy, ny, j)):
                kk = k + (di if dj == 0 else dj)
                if 0 <= kk < n:
                    A[r, idx(i + di, j + dj)] += f
                if cc and (kk < 0 or kk >= n):
                    # ghost = -2 u_0 + u_1 / 3 on the boundary face
                    A[r, r] += -2 * f
                    inner = k - (di if dj == 0 else dj)
                    A[r, idx(i - di, j - dj)] += f / 3
    return A.tocsr()


@pytest.mark.parametrize("n", [31, 32])
def test_mg_solve1_converges_to_the_discrete_solution(n):
    cfg = small_config(n=n)
    sim = orc.OracleSim(cfg)
    st = sim.store()
    rng = np.random.default_rng(5)
    L = sim.layout
    chi = rng.uniform(0.0, 3.0, (cfg.ny, cfg.nx))
    st[L["chi"], 2:-2, 2:-2] = chi
    st[L["Sy"], 2:-2, 2:-2] = rng.normal(size=(cfg.ny, cfg.nx))
    st[L["Sx"], 2:-2, 2:-2] = rng.normal(size=(cfg.ny, cfg.nx))
    sim.set_store(st)
    status, it = sim.mg_solve1(tol_rel=1e-13)
    assert status == 0 and it > 0
    out = sim.store()
    A = _operator(cfg, chi)
    for sol, rhs in (("Bx", "Sy"), ("By", "Sx")):
        want = scipy.sparse.linalg.spsolve(A, st[L[rhs], 2:-2, 2:-2].ravel()).reshape(cfg.ny, cfg.nx)
        assert rel_err(out[L[sol], 2:-2, 2:-2], want) < 1e-11, sol


# ----------------------------------------------------------------------------- push

def _push_setup(fields=None, n=500, seed=0):
    cfg = small_config(n=31)
    sim = orc.OracleSim(cfg)
    st = np.zeros_like(sim.store())
    if fields:
        for k, v in fields.items():
            st[sim.layout[k]] = v
    sim.set_store(st)
    rng = np.random.default_rng(seed)
    d = {k: np.zeros(n) for k in PLASMA_REALS}
    d["x_prev"] = rng.uniform(-3, 3, n)
    d["y_prev"] = rng.uniform(-3, 3, n)
    d["ux_half_step"] = rng.normal(0, 0.1, n)
    d["uy_half_step"] = rng.normal(0, 0.1, n)
    d["psi_half_step"] = rng.uniform(0.8, 1.2, n)
    d["w"] = np.ones(n)
    d["valid"] = np.ones(n, np.int32)
    sim.set_plasma(d)
    return cfg, sim, d


def test_push_free_streaming():
    cfg, sim, d = _push_setup()
    sim.push_plasma()
    p = sim.plasma()
    for k in ("ux", "uy", "psi"):
        assert np.array_equal(p[k], d[k + ("_half_step" if k != "psi" else "_half_step")])
    assert rel_err(p["x"], d["x_prev"] + cfg.dz * d["ux_half_step"] / d["psi_half_step"]) < 1e-15
    assert rel_err(p["y"], d["y_prev"] + cfg.dz * d["uy_half_step"] / d["psi_half_step"]) < 1e-15


def test_push_uniform_ez_is_linear_in_psi():
    """dpsi/dzeta = -q/(m c) Ez exactly for a uniform Ez (the Taylor step is exact here)."""
    Ez = 0.37
    cfg, sim, d = _push_setup({"Ez": Ez})
    sim.push_plasma()
    p = sim.plasma()
    q, m = -1.0, 1.0
    want_half = d["psi_half_step"] + cfg.dz * (-q / m * Ez)
    assert rel_err(p["psi_half_step"], want_half) < 1e-14
    assert rel_err(p["psi"], want_half + 0.5 * cfg.dz * (-q / m * Ez)) < 1e-14


def test_gather_is_exact_for_linear_psi():
    """ExmBy = -d_x Psi gathered with the nodal derivative shape is exact for linear Psi:
    a particle with u = 0 in Psi = a x + b y moves only through the resulting transverse force."""
    cfg = small_config(n=31)
    sim = orc.OracleSim(cfg)
    a, b = 0.2, -0.1
    ii = np.arange(-2, cfg.nx + 2)
    jj = np.arange(-2, cfg.ny + 2)
    X = cfg.xlo + (ii + 0.5) * cfg.dx
    Y = cfg.ylo + (jj + 0.5) * cfg.dy
    psi = a * X[None, :] + b * Y[:, None]
    st = np.zeros_like(sim.store())
    st[sim.layout["Psi"]] = psi
    sim.set_store(st)
    rng = np.random.default_rng(1)
    n = 200
    d = {k: np.zeros(n) for k in PLASMA_REALS}
    d["x_prev"] = rng.uniform(-3, 3, n)
    d["y_prev"] = rng.uniform(-3, 3, n)
    d["psi_half_step"] = np.ones(n)
    d["w"] = np.ones(n)
    d["valid"] = np.ones(n, np.int32)
    sim.set_plasma(d)
    sim.push_plasma()
    p = sim.plasma()
    # first sub-step: dux/dzeta = q/(m c) * gamma_psi * ExmBy with gamma_psi = 1, ExmBy = -a
    # after a full step with constant force F_x and psi = 1: ux_half = dz * (-a) * q/m * (1 + O(dz^2))
    approx = cfg.dz * (-1.0) * (-a)
    assert np.all(np.abs(p["ux_half_step"] / approx - 1) < 1e-2)
    approx_y = cfg.dz * (-1.0) * (-b)
    assert np.all(np.abs(p["uy_half_step"] / approx_y - 1) < 1e-2)


# ----------------------------------------------------------------------------- deposition

def test_deposit_conserves_charge_and_current():
    cfg = small_config(n=31, order=2)
    sim = orc.OracleSim(cfg)
    sim.set_store(np.zeros_like(sim.store()))
    rng = np.random.default_rng(2)
    n = 300
    d = {k: np.zeros(n) for k in PLASMA_REALS}
    d["x"] = rng.uniform(-4, 4, n)
    d["y"] = rng.uniform(-4, 4, n)
    d["w"] = rng.uniform(0.5, 1.5, n)
    d["ux"] = rng.normal(0, 0.1, n)
    d["uy"] = rng.normal(0, 0.1, n)
    d["psi"] = rng.uniform(0.9, 1.1, n)
    d["valid"] = np.ones(n, np.int32)
    sim.set_plasma(d)
    sim.deposit_plasma(["jx", "jy", -1, -1, "chi", "rhomjz"])
    st = sim.store()
    L = sim.layout
    q = -1.0
    assert math.isclose(st[L["rhomjz"]].sum(), q * d["w"].sum(), rel_tol=1e-13)
    assert math.isclose(st[L["jx"]].sum(), q * (d["w"] * d["ux"] / d["psi"]).sum(), rel_tol=1e-12)
    assert math.isclose(st[L["chi"]].sum(), (d["w"] / d["psi"]).sum(), rel_tol=1e-13)  # q^2 mu0/m = 1


def test_deposit_single_particle_brute_force():
    cfg = small_config(n=15, order=2)
    sim = orc.OracleSim(cfg)
    sim.set_store(np.zeros_like(sim.store()))
    d = {k: np.zeros(1) for k in PLASMA_REALS}
    d["x"][:] = 0.3
    d["y"][:] = -1.1
    d["w"][:] = 1.7
    d["psi"][:] = 1.0
    d["valid"] = np.ones(1, np.int32)
    sim.set_plasma(d)
    sim.deposit_plasma([-1, -1, -1, -1, -1, "rhomjz"])
    st = sim.store()[sim.layout["rhomjz"]]
    want = np.zeros_like(st)
    xm = (0.3 - (cfg.xlo + 0.5 * cfg.dx)) / cfg.dx
    ym = (-1.1 - (cfg.ylo + 0.5 * cfg.dy)) / cfg.dy
    for j in range(-2, cfg.ny + 2):
        for i in range(-2, cfg.nx + 2):
            want[j + 2, i + 2] = -1.7 * bspline(2, xm - i) * bspline(2, ym - j)
    assert rel_err(st, want) < 1e-14


# ----------------------------------------------------------------------------- beam push

def _beam_arrays(n, x, y, z, ux, uy, uz):
    f = lambda v: np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (n,)).copy())
    return {"x": f(x), "y": f(y), "z": f(z), "ux": f(ux), "uy": f(uy), "uz": f(uz), "w": np.ones(n),
            "valid": np.ones(n, np.int32), "nsub": np.zeros(n, np.int32), "charge": -1.0, "mass": 1.0}


def _beam_oracle(fields):
    cfg = small_config(n=31, nz=8)
    sim = orc.OracleSim(cfg)
    st = np.zeros_like(sim.store())
    for name, v in fields.items():
        st[sim.layout[name]] = v
    sim.set_store(st)
    return cfg, sim


def test_beam_push_free_streaming():
    """No fields: momenta are constant, x advances by dt ux / gamma, z by dt (uz / gamma - 1)."""
    cfg, sim = _beam_oracle({})
    rng = np.random.default_rng(1)
    n, s, dt, nsub = 50, 5, 2.0, 10
    z0 = cfg.zlo + (s + 0.9) * cfg.dz
    b = _beam_arrays(n, rng.uniform(-2, 2, n), rng.uniform(-2, 2, n), z0, rng.normal(0, 3, n), rng.normal(0, 3, n),
                     rng.uniform(200, 2000, n))
    b0 = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in b.items()}
    slipped = sim.push_beam_slice(b, s, 0, n, [], dt, nsub)
    g = 1.0 / np.sqrt(1 + b0["ux"] ** 2 + b0["uy"] ** 2 + b0["uz"] ** 2)
    assert slipped.size == 0
    for k in ("ux", "uy", "uz"):
        assert np.array_equal(b[k], b0[k])
    assert rel_err(b["x"] - b0["x"], dt * b0["ux"] * g) < 1e-12
    assert rel_err(b["z"] - b0["z"], dt * (b0["uz"] * g - 1.0)) < 1e-9
    assert np.all(b["nsub"] == nsub)


def test_beam_push_uniform_ez():
    """Uniform Ez, no B: uz gains dt q/m Ez per sub-cycle exactly; x follows the half-step
    velocities of the scheme."""
    E0 = 0.3
    cfg, sim = _beam_oracle({"Ez": E0})
    n, s, dt, nsub = 20, 3, 1.5, 7
    rng = np.random.default_rng(2)
    b = _beam_arrays(n, rng.uniform(-1, 1, n), 0.0, cfg.zlo + (s + 0.95) * cfg.dz, rng.normal(0, 1, n), 0.0, 50.0)
    b0 = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in b.items()}
    sim.push_beam_slice(b, s, 0, n, [], dt, nsub, do_z_push=False)
    h = dt / nsub
    qm = -1.0
    uz = b0["uz"].copy()
    x = b0["x"].copy()
    for _ in range(nsub):
        g0 = 1.0 / np.sqrt(1 + b0["ux"] ** 2 + uz ** 2)
        uz_n = uz + h * qm * E0
        g1 = 1.0 / np.sqrt(1 + b0["ux"] ** 2 + uz_n ** 2)
        x = x + h * 0.5 * b0["ux"] * g0 + h * 0.5 * b0["ux"] * g1
        uz = uz_n
    assert rel_err(b["uz"], uz) < 1e-14
    assert rel_err(b["x"] - b0["x"], x - b0["x"]) < 1e-12
    assert np.array_equal(b["z"], b0["z"])


def test_beam_push_slipped_particles_continue():
    """Slow particles fall below their slice, stop, and finish their sub-cycles in the next
    slice: the total displacement is that of n_subcycles uninterrupted sub-steps."""
    cfg, sim = _beam_oracle({})
    n, s, dt, nsub = 6, 4, 3.0, 12
    lo = cfg.zlo + s * cfg.dz
    b = _beam_arrays(n, 0.0, 0.0, lo + np.linspace(0.01, 0.2, n) * cfg.dz, 0.0, 0.0, 1.0)
    b0 = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in b.items()}
    slipped = sim.push_beam_slice(b, s, 0, n, [], dt, nsub)
    assert slipped.size == n and np.all(b["nsub"][slipped] < nsub)
    sim.push_beam_slice(b, s - 1, 0, 0, slipped, dt, nsub)
    g = 1.0 / np.sqrt(2.0)
    assert np.all(b["nsub"] == nsub)
    assert rel_err(b["z"] - b0["z"], dt * (g - 1.0)) < 1e-12


def test_beam_steps_sort_invariants():
    """Two time steps with a moving beam, then the slice sort: every valid particle lies in
    its slice's range, the order inside a slice is the previous relative order, and the
    wakefield has changed the transverse momenta."""
    from hipace2_amd.inputs import BeamSpec, make_beam
    cfg = small_config(n=15, nz=8)
    spec = BeamSpec(injection_type="fixed_weight", profile="gaussian", density=0.5,
                    position_std=(0.5, 0.5, 1.41), num_particles=400, u_mean=(0.0, 0.0, 2.0), seed=3)
    beam = make_beam(cfg, spec)
    sim = orc.OracleSim(cfg, beam)
    sim.run_steps(2, dt=3.0, n_subcycles=5)
    before = sim.beam()
    sim.sort_beam()
    b = sim.beam()
    off = b["off"]
    for s in range(cfg.nz):
        zs = b["z"][off[s]:off[s + 1]]
        assert np.all(np.floor((zs - cfg.zlo) / cfg.dz) == s)
        assert np.all(b["valid"][off[s]:off[s + 1]] == 1)
    assert np.all(b["valid"][off[-1]:] == 0)
    assert not np.array_equal(before["ux"][before["valid"] == 1], beam.ux[beam.valid == 1])
