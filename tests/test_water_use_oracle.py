"""Checks that pin the oracle's water use (subtract_use 2/3, use_alloc 0,
aggrNUsGloLakResOpt 0; routing.cpp:884-972, 1929-2137, 2193-2296, 2678-3171,
3324-3330, 3466-3512, 3590-3919, 5246-5306, 5503-5572), each on something the
use formulas do not restate: the no-use model (already pinned) at zero use,
hand-computed answers of updateNetAbstractionGW and of the year-end processing,
and the water a reservoir gives up to a use it can satisfy.

Tables are built so that the daily values are exact: v x 1e9 x days of the month
m3 per month is v km3 per day (dailyNUInit divides by 1e9 x days,
routing.cpp:966-972), with the NETABS multipliers at 1.
"""
import numpy as np
import pytest

DAYS = np.array([31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31], dtype=np.float64)


def _world(wg, seed=0):
    w = wg.World(wg.regional_spec(seed=seed))
    return w


def _monthly(daily_km3):
    """[N][12] m3 per month for per-cell daily values in km3 (same every month)."""
    d = np.asarray(daily_km3, dtype=np.float64)
    return d[:, None] * (1e9 * DAYS[None, :])


def _states_equal(a, b):
    bad = []
    for k in a:
        if isinstance(a[k], np.ndarray) and not np.array_equal(a[k], b[k]):
            bad.append(k)
    return bad


def test_zero_use_equals_the_model_without_water_use(wg, oracle):
    """With every table at 0 the water-use path is the no-use model bit for bit,
    across a year end: the use terms enter the reference's expressions as +0."""
    w = _world(wg)
    n = w.n
    f = w.forcing(seed=1, day0=0, n_days=400)
    a = oracle.Oracle(w)
    b = oracle.Oracle(w)
    b.set_water_use(np.full(n, 0.3), delayed=1)
    z = np.zeros((n, 12))
    b.water_use_year(z, z, z, z)
    for d in range(400):
        doy = d % 365 + 1
        for o in (a, b):
            o.step_day(f[d, 0], f[d, 1], f[d, 2], f[d, 3], doy)
        if doy == 365:
            b.water_use_year_end()
    assert _states_equal(a.get_state(), b.get_state()) == []
    assert not np.any(b.water_use_state())
