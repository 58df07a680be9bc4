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

from oracle import wg_oracle

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


def _humid_cells(w, k):
    a = np.asarray(w.aindex)
    ok = (a == 0) & (np.asarray(w.ldd) >= 0) & (np.asarray(w.contfreq) > 0)
    return np.flatnonzero(ok)[:k]


# (DRU, WU, CU, frgi, UI, UO, RRF) -> (UI, UO, RRF, change of the abstraction) after
# updateNetAbstractionGW, worked by hand (routing.cpp:5516-5570); km3
NAG_CASES = [
    # unsatisfied use below the irrigation share: eff 0.5, frgi 0.25, factor
    # 1 - 0.75 x 0.5 = 0.625, WUsiNew (2.5e-3 - 1.5e-3) / 0.625 = 1.6e-3, return
    # flow change 0.25 x 0.5 x (1.6e-3 - 4e-3) = -3e-4
    ((1.5e-3, 4e-3, 2e-3, 0.25, 0, 0, 0), (1.5e-3, 0, -3e-4, 3e-4)),
    # above it: WUsiNew 0, 2.5e-3 from irrigation, 2.5e-3 from other sectors,
    # return flow change 0.125 x (0 - 4e-3)
    ((5e-3, 4e-3, 2e-3, 0.25, 0, 0, 0), (2.5e-3, 2.5e-3, -5e-4, 5e-4)),
    # no irrigation withdrawal: all from other sectors, abstraction unchanged
    ((2e-3, 0, 0, 0.5, 0, 1e-3, 0), (0, 3e-3, 0, 0)),
    # delayed satisfaction covered by the other sectors' deficit
    ((-1e-3, 4e-3, 2e-3, 0.5, 0, 3e-3, 0), (0, 2e-3, 0, 0)),
    # ... beyond it: half the reduced return flow given back
    ((-3e-3, 4e-3, 2e-3, 0.5, 4e-3, 1e-3, -2e-3), (2e-3, 0, -1e-3, -1e-3)),
    # ratio below -1 clamped: all given back
    ((-3e-3, 4e-3, 2e-3, 0.5, 1e-3, 0, -2e-3), (0, 0, 0, -2e-3)),
]


def test_net_abstraction_gw_known_answers(wg):
    check_net_abstraction_gw(wg)


def check_net_abstraction_gw(wg, variant="wg"):
    """One day from set states: the unsatisfied-use sums after updateNetAbstractionGW
    match the hand-worked values, and the groundwater storage equals that of a run
    whose net abstraction from groundwater is the hand-worked adapted one."""
    w = _world(wg)
    n = w.n
    cells = _humid_cells(w, len(NAG_CASES))
    assert len(cells) == len(NAG_CASES)
    f = w.forcing(seed=2, day0=0, n_days=1)
    nug = np.full(n, 1e-3)
    frgi = np.zeros(n)
    z = np.zeros((n, 12))
    st = np.zeros((wg._abi.WG_USE_NFIELDS, n))
    nug_b = nug.copy()
    F = {name: j for j, name in enumerate(wg._abi.USE_FIELDS)}
    for c, ((dru, wu, cu, fr, ui, uo, rrf), (_, _, _, delta)) in zip(cells, NAG_CASES):
        st[F["daily_remaining"], c], st[F["withdrawal_irrig"], c], st[F["consumptive_irrig"], c] = dru, wu, cu
        st[F["unsat_irrig"], c], st[F["unsat_other"], c], st[F["reduced_return"], c] = ui, uo, rrf
        frgi[c] = fr
        nug_b[c] = nug[c] + delta
    a, b = wg_oracle.Oracle(w, variant), wg_oracle.Oracle(w, variant)
    for o, g, s in ((a, nug, st), (b, nug_b, np.zeros_like(st))):
        o.set_water_use(frgi, delayed=0)
        o.water_use_year(z, _monthly(g), z, z)
        o.set_water_use_state(s)
        o.step_day(f[0, 0], f[0, 1], f[0, 2], f[0, 3], 1)
    ua = a.water_use_state()
    for c, (_, (ui, uo, rrf, _)) in zip(cells, NAG_CASES):
        np.testing.assert_allclose([ua[F["unsat_irrig"], c], ua[F["unsat_other"], c], ua[F["reduced_return"], c]],
                                   [ui, uo, rrf], rtol=1e-12, atol=1e-18)
    ga, gb = a.get_state()["groundwater"][cells], b.get_state()["groundwater"][cells]
    np.testing.assert_allclose(ga, gb, rtol=1e-12, atol=1e-18)
    # and the abstraction matters: without the adaptation the storages differ
    assert np.abs(ga[[0, 1, 4, 5]] - oracle_gw_plain(variant, w, f, nug, frgi, cells)).min() > 1e-6


def oracle_gw_plain(variant, w, f, nug, frgi, cells):
    o = wg_oracle.Oracle(w, variant)
    z = np.zeros((w.n, 12))
    o.set_water_use(frgi, delayed=0)
    o.water_use_year(z, _monthly(nug), z, z)
    o.step_day(f[0, 0], f[0, 1], f[0, 2], f[0, 3], 1)
    return o.get_state()["groundwater"][cells][[0, 1, 4, 5]]


# (TU, UI, UO, RRF, TU_py, UI_py, UO_py, RRF_py) -> the same after the year end,
# worked by hand (routing.cpp:5257-5296)
YEAR_END_CASES = [
    ((1, 2, 3, -1, 2, 0, 0, 0), (0, 0, 0, 0, 1, 2, 3, -1)),               # less unsatisfied than last year
    ((3, 2, 1, -1, 1, 0.5, 0.25, -0.5), (2, 1.5, 0.75, -0.5, 1, 0.5, 0.25, -0.5)),
    ((3, 1, 0.5, -1, 1, 3, 0, 0), (2, 0, 0, 0, 1, 3, 0, 0)),              # sum negative: cleared
    ((3, 3, 0, -1, 1, 0, 1, 0), (2, 2, 0, -1, 1, 0, 1, 0)),               # other sectors negative
    ((3, 0, 3, -2, 1, 1, 0, 0), (2, 0, 2, 0, 1, 1, 0, 0)),                # irrigation negative
    ((3, 2, 0, 1, 1, 0, 0, 0), (2, 0, 0, 0, 1, 0, 0, 0)),                 # return flow positive
]


def test_year_end_known_answers(wg):
    check_year_end(wg)


def check_year_end(wg, variant="wg"):
    w = _world(wg)
    n = w.n
    o = wg_oracle.Oracle(w, variant)
    o.set_water_use(np.zeros(n), delayed=1)
    F = {name: j for j, name in enumerate(wg._abi.USE_FIELDS)}
    keys = ["total_unsat", "unsat_irrig", "unsat_other", "reduced_return", "unsat_prev_year",
            "unsat_irrig_prev_year", "unsat_other_prev_year", "reduced_return_prev_year"]
    st = np.zeros((wg._abi.WG_USE_NFIELDS, n))
    for c, (before, _) in enumerate(YEAR_END_CASES):
        for k, v in zip(keys, before):
            st[F[k], c] = v
    o.set_water_use_state(st)
    o.water_use_year_end()
    got = o.water_use_state()
    for c, (_, after) in enumerate(YEAR_END_CASES):
        assert [got[F[k], c] for k in keys] == list(after), c


def test_reservoir_gives_up_the_use_it_can_satisfy(wg):
    check_reservoir_use(wg)


def check_reservoir_use(wg, variant="wg"):
    """A type-2 reservoir half full: one day's use U taken from it leaves the
    storage U lower than the same day without use, the release (a rule of the
    mean outflow and the release coefficient) and everything downstream
    unchanged, and no unsatisfied use."""
    w = _world(wg)
    n = w.n
    res = np.flatnonzero((np.asarray(w.reservoir_area) > 0) & (np.asarray(w.res_type) == 2) &
                         (np.asarray(w.lake_area) == 0))
    assert len(res) > 0
    f = w.forcing(seed=2, day0=0, n_days=1)
    cap = np.asarray(w.stor_cap, dtype=np.float64)
    z = np.zeros((n, 12))

    def run(u):
        o = wg_oracle.Oracle(w, variant)
        s = o.get_state()
        s["reservoir"][res] = 0.5 * cap[res]
        o.set_state(s)
        o.set_water_use(np.zeros(n), delayed=0)
        o.water_use_year(_monthly(u), z, z, z)
        # day 2: not the first of a month, when the release coefficient of a
        # reservoir whose year starts then is set from its storage (routing.cpp:2956-2963)
        o.step_day(f[0, 0], f[0, 1], f[0, 2], f[0, 3], 2)
        return o

    # reservoirs that do not spill over their capacity on this day (a spill would
    # absorb part of the use)
    plain = run(np.zeros(n))
    res = res[plain.get_state()["reservoir"][res] < 0.9 * cap[res]]
    assert len(res) > 0
    use = np.zeros(n)
    use[res] = 0.01 * cap[res]
    runs = [plain, run(use)]
    sb, sa = runs[0].get_state(), runs[1].get_state()
    np.testing.assert_allclose(sb["reservoir"][res] - sa["reservoir"][res], use[res], rtol=1e-9)
    assert np.array_equal(runs[0].day_outputs()["discharge"], runs[1].day_outputs()["discharge"])
    assert not np.any(runs[1].water_use_state()[0][res])


def test_exhausting_use_empties_river_then_local_lake(wg):
    check_exhausting_use(wg)


def check_exhausting_use(wg, variant="wg"):
    """A use far beyond what any cell holds (1,000 and 2,000 km3/day): every river
    without a global lake or reservoir ends the day empty, its outflow the same for
    both uses (routing.cpp:3483-3492: the evaporation and the use share what the
    river can give, the transported volume is what is left of the inflow); then
    the local lake is drawn to its floor -Smax (routing.cpp:3596-3604), where its
    area reduction factor is exactly 0, the same for both uses."""
    w = _world(wg)
    n = w.n
    f = w.forcing(seed=2, day0=0, n_days=1)
    cells = np.flatnonzero((np.asarray(w.lake_area) == 0) & (np.asarray(w.reservoir_area) == 0) &
                           (np.asarray(w.ldd) >= 0) & (np.asarray(w.loc_lake) > 0))
    assert len(cells) > 1000
    z = np.zeros((n, 12))
    runs = []
    for u in (1000.0, 2000.0):
        o = wg_oracle.Oracle(w, variant)
        o.set_water_use(np.zeros(n), delayed=0)
        o.water_use_year(_monthly(np.full(n, u)), z, z, z)
        o.step_day(f[0, 0], f[0, 1], f[0, 2], f[0, 3], 2)
        runs.append((o.get_state(), o.day_outputs()["discharge"], o.water_use_state()[0]))
    (s1, q1, tu1), (s2, q2, tu2) = runs
    assert not np.any(s1["river"][cells]) and not np.any(s2["river"][cells])
    assert np.array_equal(q1[cells], q2[cells])
    assert (s1["loc_lake"][cells] < 0).all()
    assert np.array_equal(s1["loc_lake"][cells], s2["loc_lake"][cells])
    assert not np.any(s1["red_loc_lake"][cells]) and not np.any(s2["red_loc_lake"][cells])
    # what was not satisfied is kept: the larger use leaves more unsatisfied
    assert (tu2[cells] > tu1[cells]).all()


def test_river_shares_its_water_between_evaporation_and_use(wg):
    check_river_share(wg)


def check_river_share(wg, variant="wg"):
    """Cells with a river that cannot meet evaporation E plus use U give both the
    same fraction of what they can give, Max (routing.cpp:3483-3496): the use taken
    is U Max / (U + E), so 1 / taken = 1 / Max + (E / Max) / U.  Three uses per
    cell (no local lake, so the unsatisfied use is what the river left): the line
    through two predicts the third, and its slope gives E > 0 (a use served
    before the evaporation would give a slope of 0)."""
    w = _world(wg)
    n = w.n
    f = w.forcing(seed=2, day0=0, n_days=1)
    w.loc_lake[:] = 0.0  # the river's leftover is then the unsatisfied use
    # headwater cells: no upstream water, so what the river can give does not
    # change with the use of other cells
    down = np.asarray(w.downstream)
    has_up = np.zeros(n, bool)
    has_up[down[down > 0] - 1] = True
    cells = np.flatnonzero((np.asarray(w.lake_area) == 0) & (np.asarray(w.reservoir_area) == 0) &
                           (np.asarray(w.ldd) >= 0) & ~has_up)
    z = np.zeros((n, 12))

    def run(u):
        o = wg_oracle.Oracle(w, variant)
        o.set_water_use(np.zeros(n), delayed=0)
        o.water_use_year(_monthly(u), z, z, z)
        o.step_day(f[0, 0], f[0, 1], f[0, 2], f[0, 3], 2)
        return o

    base = run(np.zeros(n))
    s0 = base.get_state()
    # uses of 2, 4 and 8 times what the river held and passed on without use
    cap = s0["river"] + base.day_outputs()["discharge"]
    us = [k * np.maximum(cap, 1e-9) for k in (2.0, 4.0, 8.0)]
    taken = [u - run(u).water_use_state()[0] for u in us]
    sel = cells[(taken[0][cells] > 1e-12)]
    assert len(sel) > 100
    x = [1.0 / u[sel] for u in us]
    y = [1.0 / t[sel] for t in taken]
    slope = (y[1] - y[0]) / (x[1] - x[0])
    icpt = y[0] - slope * x[0]
    np.testing.assert_allclose(icpt + slope * x[2], y[2], rtol=1e-8)
    # E >= 0 everywhere (0 where the river has no evaporation that day), > 0 in most cells
    e = slope / icpt
    assert (e > -1e-15).all() and (e > 1e-12).mean() > 0.5


def test_global_lake_takes_its_use(wg):
    check_global_lake_use(wg)


def check_global_lake_use(wg, variant="wg"):
    """A global lake (no reservoir in the cell) below its maximum that can meet its
    use U: the day's storage drops by U (1 - e^-k) / k, the linear-reservoir solution
    dS/dt = I - PET - U - k S over one day (routing.cpp:2723-2728), or by U exactly
    where the outflow falls to 0 and the storage balance has no outflow term
    (2740-2751); nothing is left unsatisfied.  One lake at a time, so no other
    cell's water changes."""
    w = _world(wg)
    n = w.n
    f = w.forcing(seed=2, day0=0, n_days=1)
    lakes = np.flatnonzero((np.asarray(w.lake_area) > 0) & (np.asarray(w.reservoir_area) == 0))
    k_sw = np.asarray(w.params)[wg._abi.PARAM_INDEX["P_SWOUTF_C"]]
    z = np.zeros((n, 12))

    # the model starts its lakes full (setLakeWetlToMaximum); start them half full
    gmax = wg_oracle.Oracle(w, variant).get_state()["glo_lake"].copy()

    def run(u):
        o = wg_oracle.Oracle(w, variant)
        s = o.get_state()
        s["glo_lake"][lakes] = 0.5 * gmax[lakes]
        o.set_state(s)
        o.set_water_use(np.zeros(n), delayed=0)
        o.water_use_year(_monthly(u), z, z, z)
        o.step_day(f[0, 0], f[0, 1], f[0, 2], f[0, 3], 2)
        return o

    base = run(np.zeros(n)).get_state()
    checked = 0
    for c in lakes:
        if not base["glo_lake"][c] < 0.9 * gmax[c]:  # at its maximum the spill absorbs the use
            continue
        u = np.zeros(n)
        u[c] = 1e-3 * max(abs(base["glo_lake"][c]), 1e-3)
        o = run(u)
        d = base["glo_lake"][c] - o.get_state()["glo_lake"][c]
        k = k_sw[c]
        ode = u[c] * (1. - np.exp(-k)) / k
        assert np.isclose(d, ode, rtol=1e-7) or np.isclose(d, u[c], rtol=1e-7), (c, d, ode, u[c])
        assert o.water_use_state()[0][c] == 0
        checked += 1
    assert checked >= 5
