"""hpq_bin_tiles_noalloc (tile binning with its work arrays in a static device array, for
callers that cannot allocate) against hpq_bin_tiles on the same plasma, through the C ABI:
permutation and bin offsets identical (integer work), on two calls in a row with the second
one smaller, so stale entries of the shared array would show; inputs beyond the array's
limits are refused before any launch."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from helpers import PLASMA_REALS, random_plasma, small_config  # noqa: E402

pytestmark = pytest.mark.gpu


def _load(sim, d):
    import hipace2_amd._abi as abi
    n = d["x"].size
    sim.np = n
    for k in PLASMA_REALS:
        sim.plasma[k] = torch.as_tensor(d[k], dtype=torch.float64).to(sim.device)
    sim.plasma["valid"] = torch.as_tensor(d["valid"], dtype=torch.int32).to(sim.device)
    sim._pdesc = abi.Plasma(*[sim.plasma[k].data_ptr() for k in PLASMA_REALS + ["valid"]], n)


def _noalloc(sim, tile, fill=-7):
    ntx = (sim.cfg.nx + 4 + tile - 1) // tile
    nty = (sim.cfg.ny + 4 + tile - 1) // tile
    perm = torch.full((max(sim.np, 1),), fill, dtype=torch.int32, device=sim.device)
    offsets = torch.full((ntx * nty + 2,), fill, dtype=torch.int32, device=sim.device)
    sim.bin_tiles_noalloc(tile, perm, offsets)
    return perm[: sim.np], offsets


def _same(sim, tile, cfg, d):
    from oracle.oracle import tile_keys
    print(f"no-alloc binning: n={sim.np} tile={tile} grid={cfg.nx}", flush=True)
    perm_a, off_a = _noalloc(sim, tile)
    perm_b, off_b = sim.bin_tiles(tile)
    torch.cuda.synchronize()
    assert torch.equal(perm_a, perm_b)
    assert torch.equal(off_a, off_b)
    keys, nbins = tile_keys(cfg, d["x"], d["y"], d["valid"], tile)
    assert np.array_equal(perm_a.cpu().numpy(), np.argsort(keys, kind="stable"))


def test_noalloc_matches_bin_tiles_on_two_calls_second_smaller():
    from hipace2_amd.simulation import Simulation
    cfg = small_config(n=511, order=2, L=8.0)
    sim = Simulation(cfg)
    big = random_plasma(cfg, 1_000_000, seed=31)
    _load(sim, big)
    _same(sim, 32, cfg, big)
    small = random_plasma(cfg, 30_001, seed=32)
    small["valid"][::7] = 0
    _load(sim, small)
    _same(sim, 32, cfg, small)
    # and an empty plasma: offsets all zero but the last
    empty = {k: v[:0] for k, v in small.items()}
    _load(sim, empty)
    perm, off = _noalloc(sim, 32)
    torch.cuda.synchronize()
    assert perm.numel() == 0 and int(off.abs().sum()) == 0


def test_noalloc_refuses_inputs_beyond_its_limits():
    import hipace2_amd._abi as abi
    from hipace2_amd.simulation import Simulation
    cfg = small_config(n=511, order=2, L=8.0)
    sim = Simulation(cfg)
    # 2^22 + 1 particles: one more than the key array holds
    d = random_plasma(cfg, (1 << 22) + 1, seed=33)
    _load(sim, d)
    with pytest.raises(abi.HpqError, match="4194304 particles"):
        _noalloc(sim, 32)
    # 8-cell tiles on 511^2 (4226 bins) x 1024 chunks of 2^21 particles: beyond the histogram limit
    d = {k: v[: 1 << 21] for k, v in d.items()}
    _load(sim, d)
    with pytest.raises(abi.HpqError, match="chunks exceed"):
        _noalloc(sim, 8)
    # the refused calls launched nothing: a valid call right after still matches
    _same(sim, 32, cfg, d)
