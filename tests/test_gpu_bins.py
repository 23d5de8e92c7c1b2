"""hpq_bin_tiles on the GPU against the oracle's DenseBins restatement: the permutation and
the bin offsets must be identical (integer work, bit-exact)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from helpers import PLASMA_REALS, random_plasma, small_config  # noqa: E402

pytestmark = pytest.mark.gpu


def _sim_with_plasma(cfg, d):
    import hipace2_amd._abi as abi
    from hipace2_amd.simulation import Simulation
    sim = Simulation(cfg)
    n = d["x"].size
    sim.np = n
    for k in PLASMA_REALS:
        sim.plasma[k] = torch.as_tensor(d[k], dtype=torch.float64).to(sim.device)
    sim.plasma["valid"] = torch.as_tensor(d["valid"], dtype=torch.int32).to(sim.device)
    sim._pdesc = abi.Plasma(*[sim.plasma[k].data_ptr() for k in PLASMA_REALS + ["valid"]], n)
    return sim


def _check(cfg, d, tile):
    from oracle.oracle import tile_keys
    sim = _sim_with_plasma(cfg, d)
    perm, offsets = sim.bin_tiles(tile)
    keys, nbins = tile_keys(cfg, d["x"], d["y"], d["valid"], tile)
    want = np.argsort(keys, kind="stable")
    want_off = np.concatenate([[0], np.cumsum(np.bincount(keys, minlength=nbins))])
    got = perm.cpu().numpy()
    assert np.array_equal(got, want), (tile, cfg.depos_order, np.flatnonzero(got != want)[:5])
    assert np.array_equal(offsets.cpu().numpy(), want_off)


@pytest.mark.parametrize("order", [0, 1, 2, 3])
@pytest.mark.parametrize("tile", [8, 16, 32])
def test_bin_tiles_random_plasma(order, tile):
    cfg = small_config(n=63, order=order)
    _check(cfg, random_plasma(cfg, 30000, seed=order + tile), tile)


@pytest.mark.parametrize("n", [0, 1, 63, 4096, 4097])
def test_bin_tiles_small_and_ragged(n):
    cfg = small_config(n=63, order=2)
    d = random_plasma(cfg, max(n, 1), seed=n)
    if n == 0:
        d = {k: v[:0] for k, v in d.items()}
    _check(cfg, d, 16)


def test_bin_tiles_all_invalid_and_one_tile():
    cfg = small_config(n=63, order=2)
    d = random_plasma(cfg, 5000, seed=2)
    d["valid"][:] = 0
    _check(cfg, d, 16)
    d = random_plasma(cfg, 5000, seed=3)
    d["x"][:] = 0.01   # every particle in one tile: one long run of equal keys
    d["y"][:] = -0.02
    _check(cfg, d, 16)


def test_bin_tiles_large():
    """511 x 511 cells, 1M particles, 8-cell tiles: 4226 bins, 245 chunks, two scan levels."""
    cfg = small_config(n=511, order=2, L=8.0)
    _check(cfg, random_plasma(cfg, 1_000_000, seed=9), 8)
