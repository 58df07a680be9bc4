"""The calibration search (calibGammaClass::findNewGamma, calibration.cpp:266-636)
and the basin sets the batched calibration runs on.

Known-answer traces are derived by hand from calibration.cpp for annual series
whose sign and size are chosen to walk each branch; the library's search
(wg_find_new_gamma) and the oracle's restatement must both reproduce them, and
must agree call by call on searches over monotone discharge curves.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import calib_oracle as co


def lib_search():
    from watergap2_amd import _abi
    lib = _abi.load_library()
    s = _abi.GammaSearch()
    lib.wg_gamma_search_init(C.byref(s))

    def step(gamma_old, measured, simulated):
        m = np.ascontiguousarray(measured, dtype=np.float32)
        q = np.ascontiguousarray(simulated, dtype=np.float32)
        return np.float32(lib.wg_find_new_gamma(C.byref(s), float(gamma_old), len(m), _abi.ptr(m, _abi.P_f),
                                                _abi.ptr(q, _abi.P_f), None, None))
    return s, step, lib


def oracle_search():
    st = co.GammaSearchState()

    def step(gamma_old, measured, simulated):
        return co.find_new_gamma(st, gamma_old, measured, simulated)
    return st, step


OBS = np.array([10.0, 12.0, 11.0], dtype=np.float32)


def both():
    s, lstep, _ = lib_search()
    o, ostep = oracle_search()
    return (s, lstep), (o, ostep)


@pytest.mark.parametrize("which", [0, 1])
def test_kat_first_call_within_one_percent(which):
    st, step = both()[which]
    # sum of differences 0.3 against 33: 0.9 % < 1 %  -> done (calibration.cpp:327-333)
    g = step(1.0, OBS, OBS + np.float32(0.1))
    assert g == -99 and st.calib_status == 1 and st.gamma_cond == 1


@pytest.mark.parametrize("which", [0, 1])
def test_kat_bracket_then_bisect(which):
    st, step = both()[which]
    # call 1, too much runoff at gamma 1: gamma_low = 1, jump to the upper limit 5
    assert step(1.0, OBS, OBS * 1.5) == np.float32(5.0)
    assert st.gamma_low == np.float32(1.0) and st.gamma_high == np.float32(-99)
    # call 2, too little runoff at 5: gamma_high = 5, midpoint (1 + 5) / 2 = 3
    assert step(5.0, OBS, OBS * 0.5) == np.float32(3.0)
    assert st.gamma_high == np.float32(5.0)
    # call 3, too much at 3: gamma_low = 3, midpoint 4
    assert step(3.0, OBS, OBS * 1.2) == np.float32(4.0)
    # call 4, too little at 4: gamma_high = 4, midpoint 3.5
    assert step(4.0, OBS, OBS * 0.9) == np.float32(3.5)
    # call 5 within 1 %: done
    assert step(3.5, OBS, OBS) == -99 and st.calib_status == 1


@pytest.mark.parametrize("which", [0, 1])
def test_kat_lower_limit_and_correction_factor(which):
    st, step = both()[which]
    # call 1, too little runoff at gamma 1: gamma_high = 1, jump to the lower limit 0.1
    assert step(1.0, OBS, OBS * 0.5) == np.float32(0.1)
    # call 2, still too little at the limit: stays at 0.1, gamma_low = 0.1
    assert step(0.1, OBS, OBS * 0.6) == np.float32(0.1)
    assert st.gamma_low == np.float32(0.1)
    # call 3, still too little at the limit: search ends, gamma_cond 2; the mean
    # simulated discharge (0.7 Qobs) is below 0.9 Qobs, so the correction factor is
    # computed as 0.9 Qobs / Qsim = 0.9 / 0.7 (calibration.cpp:566-596)
    assert step(0.1, OBS, OBS * 0.7) == -99
    assert st.gamma_cond == 2 and st.cell_corr_factor_ind == 99 and st.calib_status == 3
    assert abs(float(st.cell_corr_factor) - 0.9 / 0.7) < 1e-6


@pytest.mark.parametrize("which", [0, 1])
def test_kat_ten_percent_rule(which):
    st, step = both()[which]
    step(1.0, OBS, OBS * 1.5)            # -> 5
    step(5.0, OBS, OBS * 1.3)            # at the limit, still too much: stays 5
    g = step(5.0, OBS, OBS * 1.05)       # still too much: ends; 1.05 < 1.1 -> 10 % rule holds
    assert g == -99 and st.gamma_cond == 2 and st.calib_status == 2 and st.cell_corr_factor == 1


@pytest.mark.parametrize("which", [0, 1])
def test_kat_no_change_terminates(which):
    st, step = both()[which]
    step(1.0, OBS, OBS * 1.5)                      # -> 5
    step(5.0, OBS, OBS * 0.5)                      # -> 3
    # the sum of differences equals the previous one: search ends (calibration.cpp:618-623)
    assert step(3.0, OBS, OBS * 0.5) == -99


def q_curve(gamma, scale, k):
    """A monotone decreasing discharge curve of gamma (like the HBV runoff)."""
    return scale / (1.0 + k * gamma)


def test_library_and_oracle_agree_on_searches():
    rng = np.random.default_rng(3)
    for trial in range(300):
        scale, k = rng.uniform(5, 500), rng.uniform(0.05, 3)
        hidden = rng.uniform(0.05, 6.0)
        n_years = int(rng.integers(1, 8))
        noise = rng.uniform(0.97, 1.03, n_years).astype(np.float32)
        obs = (q_curve(hidden, scale, k) * noise).astype(np.float32)
        if rng.uniform() < 0.2:
            obs[rng.integers(0, n_years)] = -9999.0  # a year without observations
            if (obs > -1).sum() == 0:
                continue
        (s, lstep), (o, ostep) = both()
        from watergap2_amd import _abi
        lib = _abi.load_library()
        g_l = np.float32(lib.wg_gamma_start(C.byref(s), -1.0))
        g_o = co.gamma_start(o, -1.0)
        assert g_l == g_o == np.float32(1.0)
        for it in range(40):
            sim = (q_curve(float(g_l), scale, k) * noise[::-1]).astype(np.float32)
            a = lstep(g_l, obs, sim)
            b = ostep(g_o, obs, sim)
            assert a == b, (trial, it, a, b)
            assert s.calib_status == o.calib_status and np.float32(s.cell_corr_factor) == o.cell_corr_factor
            if a < 0:
                break
            g_l = g_o = a
        else:
            pytest.fail("search did not end")


def test_basin_set_equals_separate_basins(wg, oracle):
    """Concatenated basins route exactly as the separate basins do (the basins are
    independent), checked with the CPU oracle on each."""
    from watergap2_amd.calibration import make_basins
    bs = make_basins(3, seed=5, min_cells=20, max_cells=120)
    assert bs.n == sum(w.n for w in bs.worlds)
    assert sorted(bs.rout_order.tolist()) == list(range(1, bs.n + 1))
    f = bs.forcing(seed=3, day0=0, n_days=20)
    o = oracle.Oracle(bs)
    o.set_stations(bs.basin_outlet)
    o.run(f, 1)
    whole = o.get_state()
    q_whole = o.station_discharge()
    for b, w in enumerate(bs.worlds):
        a, e = int(bs.offsets[b]), int(bs.offsets[b + 1])
        ob = oracle.Oracle(w)
        ob.set_stations(w.basin_outlet)
        ob.run(np.ascontiguousarray(f[:, :, a:e]), 1)
        sb = ob.get_state()
        for key in ("soil", "river", "groundwater", "glo_wetl", "snow"):
            assert np.array_equal(whole[key][a:e], sb[key]), (b, key)
        assert q_whole[b] == ob.station_discharge()[0]


def test_cut_basins_puts_no_station_on_an_inland_sink(wg):
    """A station at an inland sink records no discharge (routing.cpp:4232-4238), so
    cut_basins never takes one, although sinks with catchments in range exist."""
    from watergap2_amd.calibration import cut_basins, upstream_cells
    dem = wg.World(wg.global_spec(seed=3, n_land=20000, n_outlets=40, n_sinks=60))
    ldd = np.asarray(dem.ldd)
    count = upstream_cells(dem)
    sinks = np.flatnonzero(ldd == -1)
    assert ((count[sinks] >= 20) & (count[sinks] <= 2000)).sum() >= 5
    basins = cut_basins(dem, 80, np.random.default_rng(1), 20, 2000)
    assert len(basins) >= 40
    ds = dem.downstream.astype(np.int64) - 1
    for cells in basins:
        inside = np.zeros(dem.n, dtype=bool)
        inside[cells] = True
        station = [c for c in cells if ds[c] < 0 or not inside[ds[c]]]
        assert len(station) == 1
        assert ldd[station[0]] != -1


def test_thirty_year_mean_from_nine_years():
    from watergap2_amd import surrogate as sg
    # years 1-9 as run, then seven more copies of the last cycle (years 7-9)
    a = np.arange(1.0, 10.0)
    expect = (a.sum() + 7 * (7 + 8 + 9)) / 30.0
    assert sg.thirty_year_mean(a) == expect
    # a settled periodic series keeps its mean
    p = np.tile([2.0, 3.0, 7.0], 3)
    assert abs(sg.thirty_year_mean(p) - 4.0) < 1e-15


def test_runoff_target_units(wg):
    from watergap2_amd import surrogate as sg
    from watergap2_amd.calibration import make_basins
    bs = make_basins(2, seed=3, min_cells=20, max_cells=40)
    area = sg.basin_areas(bs)
    # 1 km3/yr over A km2 is 1e6/A mm/yr
    q = np.ones((2, 5))
    assert np.allclose(np.exp(sg.runoff_target(q, bs)), 1e6 / area)
    assert np.allclose(sg.runoff_target(q[1:], bs, idx=[1]), np.log(1e6 / area[1]))


def test_bracket_proposals():
    """Interpolation between the verified points that straddle the target, and
    outward steps within the search range when all points lie on one side."""
    from watergap2_amd import surrogate as sg
    # log runoff falls linearly in log gamma: y = 2 - log(g); target 2 - log(1.5)
    pts = [(1.0, 2.0), (2.0, 2.0 - np.log(2.0)), (0.5, 2.0 + np.log(2.0))]
    p = sg.bracket_proposals(pts, 2.0 - np.log(1.5))
    assert p.shape == (3,) and abs(p[0] - 1.5) < 1e-12
    assert 1.0 < p.min() and p.max() < 2.0
    # every point has too much runoff: step to larger gamma, capped at 5
    p = sg.bracket_proposals([(2.0, 3.0), (4.0, 2.5)], 1.0)
    assert np.allclose(p, [5.0, 5.0, 5.0])
    p = sg.bracket_proposals([(1.0, 3.0)], 1.0)
    assert np.allclose(p, [1.5, 2.0, 3.0])
    # too little runoff everywhere: smaller gamma, floored at 0.1
    p = sg.bracket_proposals([(0.3, 0.5), (1.0, 0.2)], 1.0)
    assert np.allclose(p, [0.2, 0.15, 0.1])
