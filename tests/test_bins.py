"""Tile binning (the CPU tiled deposition's DenseBins) in the oracle, pinned on the CPU:
the counting sort against numpy's stable argsort, the start cell against the oracle's C
shape factors, and the key formula against keys worked out by hand."""
import numpy as np
import pytest

from helpers import random_plasma, small_config


def test_dense_bins_is_stable_argsort():
    from oracle.oracle import dense_bins
    rng = np.random.default_rng(11)
    for nbins, n in [(1, 0), (1, 5), (7, 1000), (300, 20000)]:
        keys = rng.integers(0, nbins, n)
        perm, offsets = dense_bins(keys, nbins)
        assert np.array_equal(perm, np.argsort(keys, kind="stable"))
        assert np.array_equal(offsets, np.concatenate([[0], np.cumsum(np.bincount(keys, minlength=nbins))]))


@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_start_cell_matches_shape_factor_cells(order):
    from oracle.oracle import shape, start_cell
    rng = np.random.default_rng(order)
    xs = np.concatenate([rng.uniform(-3, 70, 400), np.arange(-2, 10, 0.5)])   # includes ties
    got = start_cell(order, xs)
    want = np.array([shape(order, float(x), 0)[1] for x in xs])
    assert np.array_equal(got, want)


def test_tile_keys_by_hand():
    from oracle.oracle import tile_keys
    # 12 x 12 cells on [0, 12)^2: dx = 1, cell i centred at i + 0.5; guarded box has
    # 16 cells (-2..13), so tile 8 gives 2 x 2 tiles and key = tx * 2 + ty
    cfg = small_config(n=12, order=2)
    cfg.xlo = cfg.ylo = 0.0
    cfg.xhi = cfg.yhi = 12.0
    # order 2: start cell = floor(xmid + 0.5) - 1 with xmid = x - 0.5
    x = np.array([0.5, 5.4, 5.6, 11.9, 3.0, 7.0])
    y = np.array([0.5, 0.5, 11.0, 11.9, 6.2, 6.0])
    valid = np.array([1, 1, 1, 1, 1, 0])
    # start cells cx: -1, 4, 4, 10, 2, (6); +2 -> 1, 6, 6, 12, 4 -> tiles 0, 0, 0, 1, 0
    # start cells cy: -1, -1, 10, 10, 5;   +2 -> 1, 1, 12, 12, 7 -> tiles 0, 0, 1, 1, 0
    keys, nbins = tile_keys(cfg, x, y, valid, 8)
    assert nbins == 5
    assert keys.tolist() == [0, 0, 1, 3, 0, 4]


def test_tile_keys_cover_every_bin_of_a_random_plasma():
    from oracle.oracle import dense_bins, tile_keys
    cfg = small_config(n=63, order=2)
    d = random_plasma(cfg, 50000, seed=4)
    keys, nbins = tile_keys(cfg, d["x"], d["y"], d["valid"], 16)
    assert keys.min() >= 0 and keys.max() == nbins - 1
    perm, offsets = dense_bins(keys, nbins)
    assert np.all(np.diff(keys[perm]) >= 0)
    assert offsets[-1] == d["x"].size
