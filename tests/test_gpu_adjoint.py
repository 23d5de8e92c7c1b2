"""Vector-Jacobian products of the slice operations against central finite differences of
the oracle's forward operations (CPU, FP64, deterministic sums) on the same seeded state.

For a random adjoint lam on the outputs and a random direction d of the inputs, the
product <VJP(lam), d> must equal (f(x + e d) - f(x - e d)) / 2e with f = <lam, outputs>.
Shape factors of order >= 2 are C1, so the difference quotient has an O(e^2) error; e is
chosen so that rounding (~1e-16 / e) and truncation stay below the 1e-6 bound.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from helpers import small_config  # noqa: E402
from test_gpu_kernels import _pair  # noqa: E402

pytestmark = pytest.mark.gpu

EPS = 1e-6
TOL = 1e-6


def _close(vjp, fd):
    assert abs(vjp - fd) <= TOL * max(abs(fd), 1e-12), (vjp, fd, abs(vjp - fd) / max(abs(fd), 1e-300))


def _directions(cfg, d, keys, seed):
    rng = np.random.default_rng(seed)
    scale = {"x": cfg.dx, "y": cfg.dy, "x_prev": cfg.dx, "y_prev": cfg.dy}
    return {k: rng.normal(size=d[k].shape) * scale.get(k, 0.1) for k in keys}


def _shifted(d, delta, e):
    out = {k: v.copy() for k, v in d.items()}
    for k, v in delta.items():
        out[k] = d[k] + e * v
    return out


@pytest.mark.parametrize("order", [1, 2, 3])
def test_deposit_current_vjp(order):
    cfg = small_config(order=order, dtype=2)
    sim, orc = _pair(cfg, n_particles=3000, seed=11)
    d0 = orc.plasma()
    comps = ["jx", "jy", -1, -1, "chi", "rhomjz"]
    names = [c for c in comps if c != -1]
    base = orc.store()
    for c in names:
        base[orc.layout[c]] = 0.0
    rng = np.random.default_rng(2)
    lam = {c: rng.normal(size=base.shape[1:]) for c in names}

    def f(d):
        orc.set_store(base)
        orc.set_plasma(d)
        orc.deposit_plasma(comps)
        st = orc.store()
        return sum(float((lam[c] * st[orc.layout[c]]).sum()) for c in names)

    keys = ["x", "y", "ux", "uy", "psi"]
    delta = _directions(cfg, d0, keys, 3)
    fd = (f(_shifted(d0, delta, EPS)) - f(_shifted(d0, delta, -EPS))) / (2 * EPS)
    store_adj = sim.new_store_adj()
    for c in names:
        store_adj[sim.layout[c]] = torch.as_tensor(lam[c]).to(sim.device)
    adj = sim.new_plasma_adj()
    sim.deposit_current_vjp(comps, store_adj, adj)
    vjp = sum(float((adj[k].cpu().numpy() * delta[k]).sum()) for k in keys)
    _close(vjp, fd)


@pytest.mark.parametrize("order,dtype", [(2, 2), (2, 1), (2, 0), (3, 2), (1, 2)])
def test_explicit_deposition_vjp(order, dtype):
    cfg = small_config(order=order, dtype=dtype)
    sim, orc = _pair(cfg, n_particles=3000, seed=12)
    d0 = orc.plasma()
    base = orc.store()
    for c in ("Sx", "Sy"):
        base[orc.layout[c]] = 0.0
    rng = np.random.default_rng(4)
    lam = {c: rng.normal(size=base.shape[1:]) for c in ("Sy", "Sx")}
    fields = ("Bz", "Ez", "ExmBy", "EypBx")
    dF = {c: rng.normal(size=base.shape[1:]) for c in fields}

    def f(d, e):
        st = base.copy()
        for c in fields:
            st[orc.layout[c]] += e * dF[c]
        orc.set_store(st)
        orc.set_plasma(d)
        orc.explicit_deposit()
        out = orc.store()
        return sum(float((lam[c] * out[orc.layout[c]]).sum()) for c in lam)

    keys = ["x", "y", "ux", "uy", "psi"]
    delta = _directions(cfg, d0, keys, 5)
    fd = (f(_shifted(d0, delta, EPS), EPS) - f(_shifted(d0, delta, -EPS), -EPS)) / (2 * EPS)
    sim.store.copy_(torch.as_tensor(base))
    store_adj = sim.new_store_adj()
    for c in lam:
        store_adj[sim.layout[c]] = torch.as_tensor(lam[c]).to(sim.device)
    adj = sim.new_plasma_adj()
    sim.explicit_deposition_vjp(store_adj, adj)
    sa = store_adj.cpu().numpy()
    vjp = sum(float((adj[k].cpu().numpy() * delta[k]).sum()) for k in keys)
    vjp += sum(float((sa[sim.layout[c]] * dF[c]).sum()) for c in fields)
    _close(vjp, fd)


@pytest.mark.parametrize("order,bc", [(1, "Periodic"), (2, "Periodic"), (3, "Periodic"), (2, "Reflecting"),
                                      (2, "Absorbing")])
def test_advance_plasma_particles_vjp(order, bc):
    cfg = small_config(order=order, dtype=2)
    cfg.particle_bc = bc
    sim, orc = _pair(cfg, n_particles=3000, seed=13)
    d0 = orc.plasma()
    # the random state clips some positions to exactly half a cell inside the box, a kink
    # of the order-1 nodal shape; move every position off the kinks by a small offset
    nudge = np.random.default_rng(14)
    for k, h in (("x_prev", cfg.dx), ("y_prev", cfg.dy)):
        d0[k] = d0[k] + h * nudge.uniform(0.01, 0.02, d0[k].shape)
    # a tenth of the particles close to the upper x and y boundaries, moving out of the box
    # within this push (reflected, wrapped or absorbed)
    edge = slice(0, 300)
    d0["x_prev"][edge] = cfg.xhi - nudge.uniform(0.005, 0.02, 300)
    d0["y_prev"][edge] = cfg.yhi - nudge.uniform(0.005, 0.02, 300)
    d0["ux_half_step"][edge] = nudge.uniform(0.4, 0.6, 300)
    d0["uy_half_step"][edge] = nudge.uniform(0.4, 0.6, 300)
    orc.set_plasma(d0)
    sim.set_plasma(d0)
    # fields a tenth of the unit random scale keep every particle in a smooth regime
    base = orc.store() * 0.1
    sim.store.copy_(torch.as_tensor(base))
    rng = np.random.default_rng(6)
    outs = ["x", "y", "ux_half_step", "uy_half_step", "psi_half_step", "ux", "uy", "psi"]
    lam = {k: rng.normal(size=d0["x"].shape) for k in outs}
    fields = ("Psi", "Ez", "Bx", "By", "Bz")
    dF = {c: 0.1 * rng.normal(size=base.shape[1:]) for c in fields}

    def f(d, e):
        st = base.copy()
        for c in fields:
            st[orc.layout[c]] += e * dF[c]
        orc.set_store(st)
        orc.set_plasma(d)
        orc.push_plasma(temp_slice=False)
        p = orc.plasma()
        live = p["valid"] != 0
        return sum(float((lam[k] * p[k])[live].sum()) for k in outs)

    keys = ["x_prev", "y_prev", "ux_half_step", "uy_half_step", "psi_half_step"]
    delta = _directions(cfg, d0, keys, 7)
    fd = (f(_shifted(d0, delta, EPS), EPS) - f(_shifted(d0, delta, -EPS), -EPS)) / (2 * EPS)
    out_adj = {k: torch.as_tensor(lam[k] if k in lam else np.zeros_like(lam["x"])).to(sim.device) for k in outs}
    in_adj = sim.new_plasma_adj()
    store_adj = sim.new_store_adj()
    sim.advance_plasma_particles_vjp(out_adj, in_adj, store_adj)
    sa = store_adj.cpu().numpy()
    got = {"x_prev": in_adj["x"], "y_prev": in_adj["y"], "ux_half_step": in_adj["ux_half_step"],
           "uy_half_step": in_adj["uy_half_step"], "psi_half_step": in_adj["psi_half_step"]}
    vjp = sum(float((got[k].cpu().numpy() * delta[k]).sum()) for k in keys)
    vjp += sum(float((sa[sim.layout[c]] * dF[c]).sum()) for c in fields)
    _close(vjp, fd)


# ---- the adjoint slice loop --------------------------------------------------------------

def _adjoint_sim(n=63, nz=40, peak=0.5, nbeam=20000, seed=0):
    from hipace2_amd.inputs import BeamSpec, make_beam
    from hipace2_amd.simulation import Simulation
    cfg = small_config(n=n, nz=nz, order=2, explicit=True)
    cfg.do_beam_jx_jy = False
    # the gradient is that of the converged Bx/By solve: converge the forward solves too
    cfg.mg_tol_rel = 1e-13
    spec = BeamSpec(injection_type="fixed_weight", profile="gaussian", density=peak,
                    position_mean=(0.0, 0.0, 0.0), position_std=(0.3, 0.3, 1.41),
                    num_particles=nbeam, u_mean=(0.0, 0.0, 1000.0), seed=seed)
    beam = make_beam(cfg, spec)
    return cfg, beam, Simulation(cfg, beam)


def _ez_of(sim, w):
    sim.beam["w"][: w.numel()].copy_(w)
    ez = sim.capture("Ez")
    sim.begin_step()
    sim.run_slices()
    torch.cuda.synchronize()
    out = ez.clone()
    sim.clear_captures()
    return out


def _losses(cfg):
    rng = np.random.default_rng(21)
    c = torch.as_tensor(rng.normal(size=(cfg.nz, cfg.ny, cfg.nx))).cuda()
    witness = slice(cfg.nz // 4, cfg.nz // 2)
    from hipace2_amd.adjoint import on_axis_ez_variance
    return {"linear": lambda ez: (c * ez).sum(),
            "flatness": lambda ez: on_axis_ez_variance(ez, witness) * 1e4}


@pytest.mark.parametrize("loss", ["linear", "flatness"])
def test_adjoint_loop_matches_central_differences(loss):
    from hipace2_amd.adjoint import slice_loop_ez
    cfg, beam, sim = _adjoint_sim()
    L = _losses(cfg)[loss]
    w0 = sim.beam["w"][: beam.n].clone()
    delta = w0 * torch.as_tensor(np.random.default_rng(22).normal(size=beam.n)).cuda()
    eps = 1e-3
    fd = (float(L(_ez_of(sim, w0 + eps * delta))) - float(L(_ez_of(sim, w0 - eps * delta)))) / (2 * eps)
    w = w0.clone().requires_grad_(True)
    ez = slice_loop_ez(w, sim, checkpoint_every=7)
    L(ez).backward()
    dot = float((w.grad * delta).sum())
    print(loss, "adjoint", dot, "central difference", fd)
    assert abs(dot - fd) <= 1e-4 * abs(fd), (dot, fd)


def test_adjoint_loop_checkpoint_period_invariance():
    from hipace2_amd.adjoint import slice_loop_ez
    cfg, beam, sim = _adjoint_sim(nz=24)
    L = _losses(cfg)["linear"]
    grads = []
    for k in (1, 5, 24):
        w = sim.beam["w"][: beam.n].clone().requires_grad_(True)
        L(slice_loop_ez(w, sim, checkpoint_every=k)).backward()
        grads.append(w.grad.clone())
    for g in grads[1:]:
        assert float((g - grads[0]).abs().max()) <= 1e-9 * float(grads[0].abs().max())


def test_adjoint_backward_of_a_replaced_forward_raises():
    """A backward is valid only for the last forward pass on its Simulation with the weights it
    ran with: forward(w1), forward(w2), then the first one's backward must raise, and so must a
    backward after the weights changed in place."""
    from hipace2_amd.adjoint import slice_loop_ez
    cfg, beam, sim = _adjoint_sim(nz=12, nbeam=4000)
    L = _losses(cfg)["linear"]
    w1 = sim.beam["w"][: beam.n].clone().requires_grad_(True)
    w2 = (1.1 * sim.beam["w"][: beam.n]).clone().requires_grad_(True)
    loss1 = L(slice_loop_ez(w1, sim, checkpoint_every=4))
    loss2 = L(slice_loop_ez(w2, sim, checkpoint_every=4))
    with pytest.raises(RuntimeError, match="another forward pass"):
        loss1.backward()
    sim.beam["w"][0] += 1.0
    with pytest.raises(RuntimeError, match="weights changed"):
        loss2.backward()
    sim.beam["w"][0] -= 1.0
    # a fresh forward after the failures still differentiates
    w3 = w2.detach().clone().requires_grad_(True)
    L(slice_loop_ez(w3, sim, checkpoint_every=4)).backward()
    assert torch.isfinite(w3.grad).all() and float(w3.grad.abs().max()) > 0.0


def _loading_sim(n=63, nz=40):
    from hipace2_amd import workloads  # noqa: F401
    from hipace2_amd.adjoint import BeamLoading, witness_beam
    from hipace2_amd.inputs import BeamSpec, make_beam
    from hipace2_amd.simulation import Simulation
    cfg = small_config(n=n, nz=nz, order=2, explicit=True)
    cfg.do_beam_jx_jy = False
    cfg.mg_tol_rel = 1e-13
    drive = make_beam(cfg, BeamSpec(injection_type="fixed_weight", profile="gaussian", density=0.5,
                                    position_std=(0.3, 0.3, 1.41), num_particles=20000,
                                    u_mean=(0.0, 0.0, 1000.0), seed=5))
    wit = witness_beam(cfg, 20000, -6.0, -4.0, 0.2, 0.3, 1000.0, seed=6)
    bl = BeamLoading(cfg, drive, wit, -6.0, -4.0)
    return cfg, bl, Simulation(cfg, bl.beam)


def test_energy_gain_gradient_wrt_witness_and_drive_knots():
    """The witness's energy gain (-sum over the witness slices of [Ez*jz_beam], the in-situ
    quantity of Fields.cpp:1320-1346) as a function of 8 witness knots and 8 drive knots: the
    adjoint gradient (through Ez and through jz_beam) against central differences."""
    from hipace2_amd.adjoint import energy_gain, slice_loop
    cfg, bl, sim = _loading_sim()
    wsl = bl.witness_slices()
    rng = np.random.default_rng(77)
    k0 = torch.as_tensor(rng.uniform(0.5, 1.5, 8), device="cuda")
    d0 = torch.as_tensor(rng.uniform(0.5, 1.5, 8), device="cuda")

    def plain(k, d):
        sim.beam["w"][: bl.beam.n].copy_(bl.weights(k, d))
        ez, jz = sim.capture("Ez"), sim.capture("jz_beam")
        sim.begin_step()
        sim.run_slices()
        torch.cuda.synchronize()
        out = float(energy_gain(ez, jz, wsl, cfg))
        sim.clear_captures()
        return out

    k = k0.clone().requires_grad_(True)
    d = d0.clone().requires_grad_(True)
    ez, jz = slice_loop(bl.weights(k, d), sim, checkpoint_every=8, mg_tol=1e-12)
    loss = energy_gain(ez, jz, wsl, cfg)
    loss.backward()
    assert float(loss) != 0.0
    gk, gd = k.grad.cpu().numpy(), d.grad.cpu().numpy()
    scale = max(np.abs(gk).max(), np.abs(gd).max())
    for which, g in (("witness", gk), ("drive", gd)):
        for i in (1, 4, 6):
            h = 1e-4
            e = torch.zeros(8, dtype=torch.float64, device="cuda")
            e[i] = h
            if which == "witness":
                fd = (plain(k0 + e, d0) - plain(k0 - e, d0)) / (2 * h)
            else:
                fd = (plain(k0, d0 + e) - plain(k0, d0 - e)) / (2 * h)
            print(which, i, g[i], fd)
            assert abs(g[i] - fd) <= 1e-5 * scale, (which, i, g[i], fd)
