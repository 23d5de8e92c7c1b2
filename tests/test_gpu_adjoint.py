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


@pytest.mark.parametrize("order", [1, 2, 3])
def test_advance_plasma_particles_vjp(order):
    cfg = small_config(order=order, dtype=2)
    sim, orc = _pair(cfg, n_particles=3000, seed=13)
    d0 = orc.plasma()
    # the random state clips some positions to exactly half a cell inside the box, a kink
    # of the order-1 nodal shape; move every position off the kinks by a small offset
    nudge = np.random.default_rng(14)
    for k, h in (("x_prev", cfg.dx), ("y_prev", cfg.dy)):
        d0[k] = d0[k] + h * nudge.uniform(0.01, 0.02, d0[k].shape)
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
