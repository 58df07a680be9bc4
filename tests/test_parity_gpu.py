"""GPU (HIP kernels through the C ABI) against the CPU oracle on the same inputs.

Every prognostic field after D days, the last day's per-cell discharge and the
daily station discharge, FP64 at 1e-10 relative (tests/parity.py), integers
bit-exact.
"""
import os

import numpy as np
import pytest

from parity import assert_close, compare

pytestmark = pytest.mark.gpu


def run_pair(wg, oracle, world, n_days, seed=0, first_doy=1, init_state=None, stations=None,
             profiled=False):
    f = world.forcing(seed=seed, day0=0, n_days=n_days)
    m = wg.WaterBalanceModel(world)
    o = oracle.Oracle(world)
    if init_state is not None:
        m.set_state(init_state)
        o.set_state(init_state)
    st = world.basin_outlet if stations is None else stations
    m.set_stations(st)
    o.set_stations(st)
    m.load_forcing(f)
    m.step_days(0, first_doy, n_days, profiled=profiled)
    o.run(f, first_doy)
    return m, o


def check_pair(m, o, n_days, label=""):
    assert_close(m.get_state(), o.get_state(), label=label)
    assert_close(m.day_outputs(), o.day_outputs(), label=label + "day:")
    tot, _ = m.station_discharge()
    assert_close({"discharge": tot[0]}, {"discharge": o.station_discharge()}, label=label + "station:")


def test_regional_one_year(wg, oracle, regional_world):
    print("case: regional 2000 cells, 365 days", flush=True)
    m, o = run_pair(wg, oracle, regional_world, 365)
    check_pair(m, o, 365)
    st = m.get_state()
    assert st["started"] == 1
    for k in ("soil", "canopy", "snow", "river", "groundwater"):
        assert np.isfinite(st[k]).all()


def test_regional_daily_station_series(wg, oracle, regional_world):
    print("case: regional daily station series, 40 days", flush=True)
    w = regional_world
    n_days = 40
    f = w.forcing(seed=2, day0=0, n_days=n_days)
    m = wg.WaterBalanceModel(w)
    o = oracle.Oracle(w)
    cells = np.array([w.basin_outlet[0], 5, 77, 1500], dtype=np.int32)
    m.set_stations(cells)
    o.set_stations(cells)
    m.load_forcing(f)
    m.step_days(0, 100, n_days)
    _, daily = m.station_discharge(max_days=n_days)
    prev = np.zeros(len(cells))
    for d in range(n_days):
        o.step_day(f[d, 0], f[d, 1], f[d, 2], f[d, 3], (100 - 1 + d) % 365 + 1)
        tot = o.station_discharge()
        assert_close({"discharge": daily[0, d]}, {"discharge": tot - prev}, label=f"day{d}:")
        prev = tot


def test_edge_cases(wg, oracle, regional_edge_world):
    print("case: edge-case grid 1200 cells, 120 days from day 300", flush=True)
    m, o = run_pair(wg, oracle, regional_edge_world, 120, seed=4, first_doy=300)
    check_pair(m, o, 120)
    assert (regional_edge_world.smax < 0).any()


def test_irrigation_reservoir_demand(wg, oracle):
    """Irrigation reservoirs (type 1) release against the mean demand of their cell
    and of up to five downstream cells (routing.cpp:360-403, 2972-2996), with the
    allocation coefficients of rout_prepare.cpp:895-944 stored as float like the
    reference's G_ALLOC_COEFF.5.UNF0."""
    from watergap2_amd import prepare
    print("case: edge-case grid with irrigation demand, 150 days", flush=True)
    w = wg.World(wg.regional_spec(seed=5, nrows=30, ncols=40, edge_cases=True, glolake_every=20, res_every=25))
    res = w.reservoir_area > 0
    assert (w.res_type[res] == 1).sum() >= 5
    q_typ = float(np.median(w.mean_outflow[res]))
    rng = np.random.default_rng(3)
    w.mean_nus = rng.uniform(0.0, 1.0, w.n) * q_typ * 31536000.0 / 2.0
    w.alloc_coeff = prepare.alloc_coeff(w.downstream, w.reservoir_area, w.mean_outflow).astype(np.float32)
    m, o = run_pair(wg, oracle, w, 150, seed=6, first_doy=1)
    check_pair(m, o, 150)


def test_handle_bookkeeping(wg, oracle, regional_world):
    """Rejected forcing leaves the loaded forcing in place; stations can be set
    again; a 400-day call keeps the daily series of its last block (days 365..399),
    which equals the oracle's; the profiled path takes at most 366 days."""
    print("case: handle bookkeeping, 400 days", flush=True)
    w = regional_world
    f = w.forcing(seed=9, day0=0, n_days=400)
    m = wg.WaterBalanceModel(w)
    m.load_forcing(f)
    bad = f[:3].copy()
    bad[1, 0, 17] = np.nan
    with pytest.raises(wg.WaterGapError) as e:
        m.load_forcing(bad)
    assert e.value.code == -5
    bad = f[:3].copy()
    bad[2, 1, 5] = -1.0
    with pytest.raises(wg.WaterGapError) as e:
        m.load_forcing(bad)
    assert e.value.code == -5
    m.set_stations(np.array([3, 4, 5], np.int32))
    cells = np.array([w.basin_outlet[0], 77, 1500, 77], dtype=np.int32)
    m.set_stations(cells)
    m.step_days(0, 1, 400)
    assert m.station_daily_window() == (365, 35)
    tot, daily = m.station_discharge(max_days=35)
    o = oracle.Oracle(w)
    o.set_stations(cells)
    o.run(f[:365], 1)
    prev = o.station_discharge()
    for d in range(35):
        o.step_day(f[365 + d, 0], f[365 + d, 1], f[365 + d, 2], f[365 + d, 3], d + 1)
        now = o.station_discharge()
        assert_close({"discharge": daily[0, d]}, {"discharge": now - prev}, label=f"day{365 + d}:")
        prev = now
    assert_close({"discharge": tot[0]}, {"discharge": o.station_discharge()})
    assert_close(m.get_state(), o.get_state())
    with pytest.raises(wg.WaterGapError) as e:
        m.step_days(0, 1, 400, profiled=True)
    assert e.value.code == -1


def test_random_initial_state(wg, oracle, regional_edge_world):
    """Random storages drive rarely taken branches: deep snow (threshold elevation),
    empty and negative lakes, saturated soils, land area fraction zero."""
    print("case: random initial state, 30 days", flush=True)
    w = regional_edge_world
    rng = np.random.default_rng(11)
    o0 = oracle.Oracle(w)
    st = o0.get_state()
    n = w.n
    st["canopy"] = rng.uniform(0, 1.5, n)
    st["soil"] = w.smax.astype(np.float64) * rng.uniform(0, 1.3, n)
    snow = np.zeros((n, 101))
    deep = rng.random(n) < 0.3
    snow[:, 1:] = rng.uniform(0, 50, (n, 100))
    snow[deep, 1:] = rng.uniform(500, 2500, (deep.sum(), 100))
    st["snow_elev"] = snow
    st["snow"] = snow[:, 1:].sum(axis=1) / 100.0
    for k in ("loc_lake", "glo_lake"):
        st[k] = st[k] * rng.uniform(-0.9, 1.2, n)
    for k in ("loc_wetl", "glo_wetl", "reservoir", "river", "groundwater"):
        st[k] = st[k] * rng.uniform(0, 1.2, n) + rng.uniform(0, 1e-3, n)
    st["groundwater"] = rng.uniform(0, 0.05, n)
    for k in ("red_loc_lake", "red_loc_wetl", "red_glo_lake", "red_glo_wetl", "red_res"):
        st[k] = rng.uniform(0, 1, n)
    st["lai_days"] = rng.integers(0, 45, n).astype(np.int32)
    st["lai_status"] = rng.integers(0, 2, n).astype(np.int32)
    st["lai_precsum"] = rng.uniform(0, 80, n)
    zero = rng.random(n) < 0.03
    st["laf"][zero] = 0.0
    st["laf_prev"] = st["laf"] * rng.uniform(0.9, 1.1, n)
    st["storage_transfer"] = rng.uniform(0, 5, n)
    st["started"] = 1
    m, o = run_pair(wg, oracle, w, 30, seed=8, first_doy=20, init_state=st)
    check_pair(m, o, 30)


def test_eager_equals_graph(wg, oracle, regional_world):
    print("case: eager profiled path vs graph path", flush=True)
    w = regional_world
    f = w.forcing(seed=1, day0=0, n_days=20)
    a = wg.WaterBalanceModel(w)
    b = wg.WaterBalanceModel(w)
    for mdl in (a, b):
        mdl.load_forcing(f)
    a.step_days(0, 1, 20)
    b.step_days(0, 1, 20, profiled=True)
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        if k != "started":
            assert np.array_equal(sa[k], sb[k]), k
    ms, n = b.timing()
    assert ms[0] > 0 and ms[1] > 0 and n[0] == 20


def test_level_schedule_invariance(wg, regional_world):
    """Results do not depend on how basins are packed into routing groups, how the
    groups are placed over the XCDs, or on the chain workgroup size (bit-identical)."""
    print("case: schedule invariance", flush=True)
    w = regional_world
    f = w.forcing(seed=3, day0=0, n_days=30)
    states = []
    for cells, threads, xcds in ((1024, 256, 8), (1, 64, 3), (100000, 128, 1)):
        os.environ["WG_GROUP_CELLS"] = str(cells)
        os.environ["WG_CHAIN_THREADS"] = str(threads)
        os.environ["WG_XCDS"] = str(xcds)
        try:
            m = wg.WaterBalanceModel(w)
        finally:
            del os.environ["WG_GROUP_CELLS"]
            del os.environ["WG_CHAIN_THREADS"]
            del os.environ["WG_XCDS"]
        m.load_forcing(f)
        m.step_days(0, 50, 30)
        states.append(m.get_state())
    for s in states[1:]:
        for k in s:
            if k != "started":
                assert np.array_equal(s[k], states[0][k]), k


def test_members_params_and_perturbation(wg, oracle, regional_world):
    """Members with their own gamma/CFA/Smax and perturbed forcing each match a
    separate oracle run with the same inputs."""
    print("case: 3 members, perturbed forcing and parameters, 45 days", flush=True)
    w = regional_world
    n, M, n_days = w.n, 3, 45
    rng = np.random.default_rng(5)
    gamma = np.tile(w.params[0], (M, 1)) * rng.uniform(0.8, 1.2, (M, n))
    cfa = rng.uniform(0.9, 1.1, (M, n))
    smax = (np.tile(w.smax, (M, 1)) * rng.uniform(0.8, 1.2, (M, n))).astype(np.float32)
    seeds = np.array([11, 12, 13], dtype=np.uint64)
    f = w.forcing(seed=0, day0=0, n_days=n_days)
    refs = []
    for k in range(M):
        o = oracle.Oracle(w)
        o.set_params(gamma[k], cfa[k], smax[k])
        o.set_stations(w.basin_outlet)
        for d in range(n_days):
            t, p = oracle.perturb_forcing(int(seeds[k]), d, f[d, 0], f[d, 1], 0.2, 1.0)
            o.step_day(t, p, f[d, 2], f[d, 3], d + 1)
        refs.append((o.get_state(), o.station_discharge()))
    # both cell-kernel builds: snow chunks of 5 pairs (held to three wavefronts per
    # SIMD) and of 10
    for chunk in (5, 10):
        m = wg.WaterBalanceModel(w, n_members=M)
        m.set_launch_config(256, 256, chunk, 3)
        m.set_member_params(gamma, cfa, smax)
        m.set_forcing_perturbation(seeds, 0.2, 1.0)
        m.set_stations(w.basin_outlet)
        m.load_forcing(f)
        m.step_days(0, 1, n_days)
        tot, _ = m.station_discharge()
        for k in range(M):
            assert_close(m.get_state(k), refs[k][0], label=f"chunk {chunk} m{k}:")
            assert_close({"discharge": tot[k]}, {"discharge": refs[k][1]}, label=f"chunk {chunk} m{k}:")


def test_global_grid_20_days(wg, oracle, global_world):
    print("case: global 67420 cells, 20 days", flush=True)
    m, o = run_pair(wg, oracle, global_world, 20, seed=1, first_doy=160)
    check_pair(m, o, 20)
    info = m.schedule_info()
    assert info["groups"] >= 1 and info["max_depth"] == global_world.max_level + 1


def test_launch_configs_and_autotune_are_bit_identical(wg, regional_world):
    """Every launch configuration the autotuner chooses from gives the same bits,
    and the autotuner leaves the model state as it found it."""
    print("case: launch configurations and autotune, regional 30 days", flush=True)
    w = regional_world
    f = w.forcing(seed=8, day0=0, n_days=30)
    ref = None
    for cell, chain, chunk, mode in ((256, 256, 5, 0), (128, 64, 10, 0), (192, 128, 5, 0), (64, 256, 10, 0),
                                     (256, 256, 5, 1), (128, 64, 10, 1), (256, 128, 5, 1),
                                     (256, 256, 5, 2), (128, 64, 10, 2), (192, 128, 5, 2),
                                     (256, 256, 5, 3), (128, 512, 10, 3)):
        m = wg.WaterBalanceModel(w)
        m.load_forcing(f)
        m.set_launch_config(cell, chain, chunk, mode)
        m.step_days(0, 1, 30)
        st = m.get_state()
        if ref is None:
            ref = st
        else:
            for k in st:
                if k != "started":
                    assert np.array_equal(st[k], ref[k]), (cell, chain, chunk, mode, k)
    m = wg.WaterBalanceModel(w)
    m.load_forcing(f)
    m.step_days(0, 1, 10)
    before = m.get_state()
    cfg, ms = m.autotune(n_days=5)
    assert cfg[0] in (128, 256) and cfg[1] in (64, 128, 256, 512) and cfg[2] in (5, 10) and cfg[3] in (0, 1, 2, 3) and ms > 0
    after = m.get_state()
    for k in before:
        assert np.array_equal(np.asarray(before[k]), np.asarray(after[k])), k
    m.step_days(10, 11, 20)
    st = m.get_state()
    for k in st:
        if k != "started":
            assert np.array_equal(st[k], ref[k]), ("after autotune", k)


def test_vertical_closure_on_gpu(wg):
    """Per-cell water balance of the GPU day loop, 120 days, CFA = 1 and no built-up
    area: storage change after the land-area rescaling equals precipitation minus
    evapotranspiration, surface runoff and recharge to < 1e-9 mm/day."""
    print("case: GPU closure, regional 2000 cells, 120 days", flush=True)
    w = wg.World(wg.regional_spec(seed=0))
    w.built_up[:] = 0.0
    n = w.n
    f = w.forcing(seed=4, day0=0, n_days=120)
    m = wg.WaterBalanceModel(w)
    m.set_member_params(cfa=np.ones((1, n)))
    m.load_forcing(f)
    smax = w.smax.astype(np.float64)

    def storage(st):
        return st["canopy"] + st["snow_elev"][:, 1:].sum(axis=1) / 100.0 + st["soil"]

    worst = 0.0
    for d in range(120):
        before = m.get_state()
        m.step_days(d, d + 1, 1)
        after = m.get_state()
        out = m.day_outputs()
        laf = before["laf"]
        laf_prev = before["laf_prev"] if before["started"] else laf
        ok = (laf > 0) & (smax > 0)
        flux = out["prec"] - out["cell_aet_uncorr"] - out["surface_runoff"] - out["gw_recharge"]
        res = storage(after)[ok] - storage(before)[ok] * laf_prev[ok] / laf[ok] - flux[ok]
        worst = max(worst, float(np.max(np.abs(res))))
    assert worst < 1e-9, worst


def test_bit_identical_to_oracle(wg, oracle, regional_edge_world):
    """Beyond the 1e-10 bar: with the shared exp/log/pow and the reference's
    operation order the GPU reproduces the oracle's bits in every field."""
    print("case: bit identity, edge world 120 days", flush=True)
    m, o = run_pair(wg, oracle, regional_edge_world, 120, seed=6)
    for a, b in ((m.get_state(), o.get_state()), (m.day_outputs(), o.day_outputs())):
        diff = [k for k in b if k != "started" and not np.array_equal(np.asarray(a[k]), np.asarray(b[k]))]
        assert not diff, diff
    tot, _ = m.station_discharge()
    assert np.array_equal(tot[0], o.station_discharge())
