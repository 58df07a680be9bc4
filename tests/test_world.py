"""Synthetic grids: sizes, types, determinism and hydrological consistency."""
import numpy as np


def test_regional_grid(regional_world):
    w = regional_world
    assert w.n == 2000
    assert w.n_basins == 1  # single outlet
    assert (w.downstream == 0).sum() == 1
    assert w.elevation.shape == (2000, 101)
    assert (w.elevation[:, 1:] >= 1).all()
    assert (np.diff(w.elevation[:, 1:].astype(int), axis=1) >= 0).all()  # sub-cell 1 lowest
    assert set(np.unique(w.aindex)) <= {0, 1}
    assert w.smax.dtype == np.float32 and w.rgmax.dtype == np.int16
    assert ((w.smax >= 50) & (w.smax <= 500)).all()
    g = w.params[0]
    assert ((g >= 0.1) & (g <= 5.0)).all()
    assert (w.reservoir_area > 0).sum() >= 5


def test_global_grid(global_world):
    w = global_world
    assert w.n == 67420
    sizes = np.bincount(w.basin)
    assert 500 <= w.n_basins <= 1500
    assert (w.ldd == -1).sum() >= 1  # inland sinks
    assert w.max_level >= 100  # 104 routing levels
    assert sizes.max() > 2000


def test_deep_global_grid(wg):
    """The same land mask drained by 60 ocean outlets: a routing network of the
    depth the design plans for (about 350-400 levels)."""
    w = wg.World(wg.global_spec(seed=1, n_outlets=60))
    sizes = np.bincount(w.basin)
    assert w.n == 67420
    assert w.max_level + 1 >= 350
    assert sizes.max() > 10000


def test_deterministic(wg):
    a = wg.World(wg.regional_spec(seed=3, nrows=10, ncols=12))
    b = wg.World(wg.regional_spec(seed=3, nrows=10, ncols=12))
    assert np.array_equal(a.altitude, b.altitude) and np.array_equal(a.rout_order, b.rout_order)
    fa = a.forcing(seed=9, day0=10, n_days=5)
    fb = b.forcing(seed=9, day0=0, n_days=15)[10:]
    assert np.array_equal(fa, fb)  # counter-based: any day range regenerates identically


def test_forcing_distributions(regional_world):
    f = regional_world.forcing(seed=0, day0=0, n_days=365)
    wet = f[:, 1] > 0
    assert 0.35 < wet.mean() < 0.45
    assert abs(f[:, 1][wet].mean() - 6.4) < 0.5  # gamma(0.8, 8) mean
    assert f[:, 2].min() >= 50 and f[:, 2].max() <= 300
    assert f[:, 3].min() >= 250 and f[:, 3].max() <= 400
    tmean = f[:, 0].mean(axis=1)
    assert tmean[200] > tmean[15] + 15  # seasonal cycle peaking in summer
