"""Water use on the GPU (wg_set_water_use: subtract_use 2, use_alloc 0,
aggrNUsGloLakResOpt 0) against the oracle, bit for bit: two model years of
seeded tables with return flows (negative net use) and cells whose use exceeds
what the river holds, so that the local lake, the reservoir's 10 % floor and the
unsatisfied-use bookkeeping all act; delayed satisfaction on and off; every chain
kernel; two members with their own parameters; a call boundary at the year end
and the year end inside a call."""
import numpy as np
import pytest

from parity import assert_close

pytestmark = pytest.mark.gpu

DAYS = np.array([31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31], dtype=np.float64)


def use_tables(w, seed, years=2):
    """[years][N][12] m3 per month: net use from surface water (a tenth of the cells
    twenty times larger, some negative), from groundwater, irrigation withdrawal
    and consumptive use; and G_FRACTRETURNGW_IRRIG."""
    rng = np.random.default_rng(seed)
    n = w.n
    scale = np.asarray(w.cell_area, dtype=np.float64) * 2e-6  # 2 mm over the cell, km3
    nus = rng.uniform(-0.3, 1.5, (years, n, 12)) * scale[None, :, None]
    nus[:, rng.random(n) < 0.1, :] *= 20.0
    nug = rng.uniform(0.0, 0.5, (years, n, 12)) * scale[None, :, None]
    wu = rng.uniform(0.0, 2.0, (years, n, 12)) * scale[None, :, None]
    cu = wu * rng.uniform(0.2, 0.9, (years, n, 12))
    m3 = 1e9 * DAYS[None, None, :]
    return [t * m3 for t in (nus, nug, wu, cu)], rng.uniform(0.0, 1.0, n)


def oracle_run(oracle, w, f, tabs, frgi, delayed, n_days, params=None):
    o = oracle.Oracle(w)
    if params is not None:
        o.set_params(*params)
    o.set_water_use(frgi, delayed)
    o.water_use_year(*[t[0] for t in tabs])
    for d in range(n_days):
        doy = d % 365 + 1
        o.step_day(f[d, 0], f[d, 1], f[d, 2], f[d, 3], doy)
        if doy == 365:
            o.water_use_year_end()
            o.water_use_year(*[t[(d + 1) // 365] for t in tabs])
    return o


@pytest.mark.parametrize("delayed", [1, 0])
@pytest.mark.parametrize("mode", [3, 2, 1, 0])
def test_water_use_matches_oracle(wg, oracle, regional_edge_world, delayed, mode):
    w = regional_edge_world
    M, n_days = 2, 400
    print(f"case: water use, edge world, delayed {delayed}, chain mode {mode}, {M} members, {n_days} days",
          flush=True)
    tabs, frgi = use_tables(w, seed=11)
    rng = np.random.default_rng(5)
    gamma = np.tile(w.params[0], (M, 1)) * rng.uniform(0.8, 1.2, (M, w.n))
    cfa = rng.uniform(0.8, 1.2, (M, w.n))
    smax = np.tile(w.smax, (M, 1)).astype(np.float32)
    f = w.forcing(seed=3, day0=0, n_days=n_days)
    m = wg.WaterBalanceModel(w, n_members=M, options={"subtract_use": 2})
    m.set_launch_config(256, 512 if mode == 3 else 256, 10, mode)
    m.set_member_params(gamma, cfa, smax)
    m.set_water_use(*tabs, frgi, year0_day=0, delayed=delayed)
    m.load_forcing(f)
    # a call that ends on day 365 and one with the year end inside
    bounds = (0, 365, n_days) if mode in (3, 1) else (0, 300, n_days)
    for a, b in zip(bounds[:-1], bounds[1:]):
        m.step_days(a, a % 365 + 1, b - a)
    for k in range(M):
        o = oracle_run(oracle, w, f, tabs, frgi, delayed, n_days, (gamma[k], cfa[k], smax[k]))
        gs, os_ = m.get_state(k), o.get_state()
        assert_close(gs, os_, label=f"delayed {delayed} mode {mode} m{k}:")
        diff = [key for key in os_ if key != "started" and not np.array_equal(np.asarray(gs[key]), os_[key])]
        assert diff == [], diff
        gu, ou = m.water_use_state(k), o.water_use_state()
        bad = [name for j, name in enumerate(wg._abi.USE_FIELDS) if not np.array_equal(gu[j], ou[j])]
        assert bad == [], bad
    # the use acted: unsatisfied use, adapted return flows, some local lake drawn below zero
    assert np.abs(ou[0]).max() > 0 and np.abs(ou[5]).max() > 0
    assert (o.get_state()["loc_lake"] < 0).any()


def test_zero_tables_equal_no_water_use(wg, regional_edge_world):
    """All tables 0: the water-use handle's state equals a plain handle's."""
    print("case: water use with zero tables vs subtract_use 0, 120 days", flush=True)
    w = regional_edge_world
    f = w.forcing(seed=4, day0=0, n_days=120)
    a = wg.WaterBalanceModel(w)
    b = wg.WaterBalanceModel(w, options={"subtract_use": 3})
    z = np.zeros((1, w.n, 12))
    b.set_water_use(z, z, z, z, np.full(w.n, 0.5), delayed=1)
    for m in (a, b):
        m.load_forcing(f)
        m.step_days(0, 1, 120)
    sa, sb = a.get_state(), b.get_state()
    for key in sa:
        assert np.array_equal(np.asarray(sa[key]), np.asarray(sb[key])), key
    assert not np.any(b.water_use_state())


def test_refused_calls(wg, regional_edge_world):
    """Tables on a handle without water use, days outside the tables, a step with
    no tables and a step after wg_annual_init without new tables are refused."""
    w = regional_edge_world
    f = w.forcing(seed=4, day0=0, n_days=40)
    z = np.zeros((1, w.n, 12))
    plain = wg.WaterBalanceModel(w)
    with pytest.raises(RuntimeError):
        plain.set_water_use(z, z, z, z, np.zeros(w.n))
    m = wg.WaterBalanceModel(w, options={"subtract_use": 2})
    m.load_forcing(f)
    with pytest.raises(RuntimeError):
        m.step_days(0, 1, 10)
    m.set_water_use(z, z, z, z, np.zeros(w.n), year0_day=20)
    with pytest.raises(RuntimeError):
        m.step_days(0, 1, 10)
    m.step_days(20, 21, 10)
    m.annual_init(w)
    with pytest.raises(RuntimeError):
        m.step_days(30, 31, 5)
    m.set_water_use(z, z, z, z, np.zeros(w.n), year0_day=20)
    m.step_days(30, 31, 5)


def test_water_use_global_grid_two_partitions(wg, oracle, global_world):
    """The global grid (67,420 cells) with water use, its deepest routing groups
    run as a second partition beside the others: 380 days across a year end,
    every cell and every water-use field against the oracle."""
    w = global_world
    n_days = 380
    print(f"case: water use, global grid, two partitions, {n_days} days", flush=True)
    tabs, frgi = use_tables(w, seed=13)
    f = w.forcing(seed=1, day0=0, n_days=n_days)
    m = wg.WaterBalanceModel(w, options={"subtract_use": 2})
    assert len(m.partitions()) == 2
    m.set_water_use(*tabs, frgi, year0_day=0, delayed=1)
    m.load_forcing(f)
    m.step_days(0, 1, 200)
    m.step_days(200, 201, n_days - 200)
    o = oracle_run(oracle, w, f, tabs, frgi, 1, n_days)
    gs, os_ = m.get_state(), o.get_state()
    diff = [key for key in os_ if key != "started" and not np.array_equal(np.asarray(gs[key]), os_[key])]
    assert diff == [], diff
    assert np.array_equal(m.water_use_state(), o.water_use_state())
