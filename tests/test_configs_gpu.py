"""The benchmark configurations at their own sizes on the GPU.

* Global grid (67,420 cells, seed 1): one year against the oracle on every cell,
  and ten years against the oracle on whole basins cut out of it (basins are
  independent, so the cut-out cells must follow the same trajectories).
* Deep network (the same mask with 60 ocean outlets, 352 routing levels, the
  bench's deep leg): 60 days against the oracle on every cell.
* Ensemble: 256 perturbed members of the global grid for a year; three members
  against oracle runs of their inputs for 30 days, a batched member with the
  perturbation off against the one-member run of its parameters, and the year's
  storages finite and non-negative.
* Calibration: 120 basins (log-uniform 50-5,000 cells) cut from one seeded global
  DEM, the reference's gamma search for 30 years with its steps replayed by the
  oracle's restatement of findNewGamma on the GPU's discharges, 64 (gamma, CFA)
  candidates per basin in one run, and the largest basin's candidates against
  the oracle for 3 years.

Tolerances are those of tests/parity.py (1e-10 relative, integers exact).
"""
import os

import numpy as np
import pytest

from parity import assert_close

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def oracle_for(oracle, world):
    o = oracle.Oracle(world)
    o.set_threads(THREADS)
    o.set_stations(world.basin_outlet)
    return o


def test_global_grid_one_year(wg, oracle, global_world):
    w = global_world
    print("case: global 67420 cells, 365 days, every cell against the oracle", flush=True)
    f = w.forcing(seed=1, day0=0, n_days=365)
    m = wg.WaterBalanceModel(w)
    m.set_stations(w.basin_outlet)
    m.load_forcing(f)
    m.step_days(0, 1, 365)
    o = oracle_for(oracle, w)
    o.run(f, 1)
    assert_close(m.get_state(), o.get_state(), label="year 1:")
    assert_close(m.day_outputs(), o.day_outputs(), label="year 1 day:")
    tot, daily = m.station_discharge(max_days=365)
    assert_close({"discharge": tot[0]}, {"discharge": o.station_discharge()}, label="year 1 station:")
    assert np.isclose(daily[0].sum(axis=0), tot[0], rtol=1e-12, atol=1e-15).all()
    m.close()


def whole_basin_cells(w, budget):
    """The deepest basin and then basins with a global lake or reservoir, whole,
    up to about `budget` cells (ascending cell indices)."""
    sizes = np.bincount(w.basin, minlength=w.n_basins)
    chosen = [int(w.basin[int(np.argmax(w.level))])]
    total = int(sizes[chosen[0]])
    special = np.unique(w.basin[(w.glo_lake > 0) | (w.reservoir_area > 0)])
    for b in special[np.argsort(sizes[special])]:
        if int(b) in chosen or total + sizes[b] > budget:
            continue
        chosen.append(int(b))
        total += int(sizes[b])
    return np.flatnonzero(np.isin(w.basin, chosen)).astype(np.int32)


def test_global_grid_ten_years_on_whole_basins(wg, oracle, global_world):
    w = global_world
    cells = whole_basin_cells(w, 12000)
    sub = w.subset(cells)
    assert sub.n_basins >= 2 and (sub.reservoir_area > 0).any() and (sub.glo_lake > 0).any()
    assert sub.max_level == w.max_level
    print(f"case: global 67420 cells, 10 years; oracle on {sub.n} cells of {sub.n_basins} whole basins", flush=True)
    m = wg.WaterBalanceModel(w)
    m.set_stations(w.basin_outlet)
    o = oracle_for(oracle, sub)
    for y in range(10):
        f = w.forcing(seed=1, day0=365 * y, n_days=365)
        m.load_forcing(f)
        m.step_days(0, 1, 365)
        o.run(np.ascontiguousarray(f[:, :, cells]), 1)
    g = m.get_state()
    ref = o.get_state()
    got = {k: (np.asarray(v)[cells] if k != "started" else v) for k, v in g.items()}
    assert_close(got, ref, label="year 10:")
    assert g["started"] == 1
    m.close()


def test_deep_network_60_days(wg, oracle):
    w = wg.World(wg.global_spec(seed=1, n_outlets=60))
    print(f"case: deep network, {w.n} cells, {w.max_level + 1} levels, 60 days against the oracle", flush=True)
    assert w.max_level + 1 >= 300
    n_days = 60
    f = w.forcing(seed=3, day0=0, n_days=n_days)
    m = wg.WaterBalanceModel(w)
    m.set_stations(w.basin_outlet)
    m.load_forcing(f)
    m.step_days(0, 100, n_days)
    o = oracle_for(oracle, w)
    o.run(f, 100)
    assert_close(m.get_state(), o.get_state(), label="deep:")
    assert_close(m.day_outputs(), o.day_outputs(), label="deep day:")
    tot, _ = m.station_discharge()
    assert_close({"discharge": tot[0]}, {"discharge": o.station_discharge()}, label="deep station:")
    assert m.schedule_info()["max_depth"] == w.max_level + 1


def test_ensemble_256_members_one_year(wg, oracle, global_world):
    w = global_world
    n, M = w.n, 256
    print("case: ensemble 256 members x 67420 cells, 365 days", flush=True)
    rng = np.random.default_rng(2024)
    gamma = np.tile(w.params[wg.PARAM_INDEX["P_GAMRUN_C"]], (M, 1)) * rng.uniform(0.8, 1.2, (M, n))
    smax = (np.tile(w.smax, (M, 1)) * rng.uniform(0.8, 1.2, (M, n))).astype(np.float32)
    seeds = np.arange(1000, 1000 + M, dtype=np.uint64)
    f = w.forcing(seed=1, day0=0, n_days=365)
    m = wg.WaterBalanceModel(w, n_members=M)
    m.set_member_params(gamma=gamma, smax=smax)
    m.set_forcing_perturbation(seeds, 0.2, 1.0)
    m.set_stations(w.basin_outlet)
    m.load_forcing(f)
    m.step_days(0, 1, 30)
    tot30, _ = m.station_discharge()
    cfa = w.params[wg.PARAM_INDEX["P_CFA"]]
    for k in (0, 127, 255):
        o = oracle_for(oracle, w)
        o.set_params(gamma[k], cfa, smax[k])
        for d in range(30):
            t, p = oracle.perturb_forcing(int(seeds[k]), d, f[d, 0], f[d, 1], 0.2, 1.0)
            o.step_day(t, p, f[d, 2], f[d, 3], d + 1)
        assert_close(m.get_state(k), o.get_state(), label=f"member {k} day 30:")
        assert_close({"discharge": tot30[k]}, {"discharge": o.station_discharge()}, label=f"member {k} station:")
    m.step_days(30, 31, 335)
    tot, _ = m.station_discharge()
    assert np.isfinite(tot).all() and (tot >= 0).all()
    for k in (0, 63, 127, 191, 255):
        st = m.get_state(k)
        for name in ("canopy", "soil", "snow", "groundwater", "river", "loc_lake", "glo_lake", "reservoir"):
            a = np.asarray(st[name])
            assert np.isfinite(a).all(), (k, name)
        for name in ("canopy", "soil", "snow", "river"):
            assert (np.asarray(st[name]) >= 0).all(), (k, name)
    # members differ (the perturbation and parameters act), but not wildly
    spread = tot.std(axis=0) / np.maximum(tot.mean(axis=0), 1e-30)
    big = tot.mean(axis=0) > 1e-3
    assert (spread[big] > 0).all() and (spread[big] < 1.0).all()
    m.close()
    # a batched member without perturbation gives the one-member run of its parameters
    e = wg.WaterBalanceModel(w, n_members=M)
    e.set_member_params(gamma=gamma, smax=smax)
    e.load_forcing(f)
    e.step_days(0, 1, 30)
    single = wg.WaterBalanceModel(w)
    single.set_member_params(gamma=gamma[127:128], smax=smax[127:128])
    single.load_forcing(f)
    single.step_days(0, 1, 30)
    a, b = e.get_state(127), single.get_state()
    for name in a:
        if name != "started":
            assert np.array_equal(np.asarray(a[name]), np.asarray(b[name])), name
    e.close()
    single.close()


class RecordingRunner:
    """A BasinRunner whose annual discharges are kept, run by run."""

    def __init__(self, runner):
        self.r = runner
        self.n_members = runner.n_members
        self.basins = runner.basins
        self.runs = []

    def run(self, gamma, cfa=None, n_years=30):
        q = self.r.run(gamma, cfa, n_years=n_years)
        self.runs.append(q[0].copy())
        return q


def test_calibration_120_basins_30_years(wg, oracle):
    from oracle import calib_oracle as co
    from watergap2_amd.calibration import BasinRunner, make_basins, plain_search
    bs = make_basins(120, seed=2024)
    sizes = np.diff(bs.offsets)
    print(f"case: calibration, 120 basins cut from one DEM ({bs.n} cells, {sizes.min()}-{sizes.max()}), "
          f"30 years", flush=True)
    assert bs.n_basins == 120 and sizes.min() >= 50 and sizes.max() <= 5000
    years = 30
    hidden = np.random.default_rng(99).uniform(0.1, 5.0, bs.n_basins)
    runner = BasinRunner(bs, n_members=1, forcing_years=3)
    measured = runner.run(hidden[None, :], n_years=years)[0].astype(np.float32)
    rec = RecordingRunner(runner)
    res = plain_search(rec, measured, n_years=years)
    runner.close()
    assert res["unfinished"] == 0
    assert (res["status"] >= 1).all()
    # replay every basin's search with the oracle's findNewGamma on the GPU's discharges
    for i in range(bs.n_basins):
        st = co.GammaSearchState()
        g = co.gamma_start(st, -1.0)
        traj = [float(g)]
        for run in rec.runs[:res["runs"][i]]:
            g_new = co.find_new_gamma(st, g, measured[i], run[i].astype(np.float32))
            if g_new < 0:
                break
            g = g_new
            traj.append(float(g))
        assert traj == res["trajectories"][i], (i, traj, res["trajectories"][i])
    # 64 candidates per basin in one run: 8 gammas x 8 CFAs
    gv = np.linspace(0.1, 5.0, 8)
    cv = np.linspace(0.5, 1.5, 8)
    gam = np.repeat(gv, 8)[:, None] * np.ones((1, bs.n_basins))
    cfa = np.tile(cv, 8)[:, None] * np.ones((1, bs.n_basins))
    K = gam.shape[0]
    cand = BasinRunner(bs, n_members=K, forcing_years=3)
    q = cand.run(gam, cfa, n_years=years)
    cand.close()
    assert q.shape == (K, bs.n_basins, years)
    # With CFA > 1 the reference's land-area rescaling can leave a soil storage a
    # hair below zero, whose pow() is NaN (routing.cpp:5343-5352, daily.cpp:1126);
    # the oracle does the same.  Every candidate with CFA <= 1 stays finite.
    finite = np.isfinite(q).all(axis=2)  # [candidate][basin]
    print(f"non-finite candidate series: {int((~finite).sum())} of {finite.size}", flush=True)
    assert finite[cfa[:, 0] <= 1.0].all()
    assert (~finite).mean() < 0.05
    assert (q[np.isfinite(q)] >= 0).all()
    # at a fixed CFA the long-term discharge falls as gamma rises: the runoff
    # fraction (S/Smax)^gamma falls (daily.cpp:1126).  (CFA scales the land runoff
    # up but also the open-water balance P - PET, daily.cpp:773, 1155, so its sign
    # depends on the basin.)
    mean = q.mean(axis=2).reshape(8, 8, bs.n_basins)  # [gamma][cfa][basin]
    with np.errstate(invalid="ignore"):
        dg = np.diff(mean, axis=0)
        okg = np.isfinite(dg)
        falls_g = (dg[okg] <= 1e-12 * mean[:-1][okg]).mean()
    print(f"discharge falls with gamma in {falls_g:.4f} of the steps", flush=True)
    assert falls_g > 0.95
    # the largest basin's candidates against the oracle for 3 years
    big = int(np.argmax(sizes))
    one = make_one(bs, big)
    r1 = BasinRunner(one, n_members=2, forcing_years=3, forcing=np.ascontiguousarray(cand_forcing(bs, big)))
    q1 = r1.run(gam[[9, 50], big:big + 1], cfa[[9, 50], big:big + 1], n_years=3)
    assert np.array_equal(q1[:, 0, :], q[[9, 50], big, :3])  # basins run independently
    st = [r1.model.get_state(k) for k in range(2)]
    r1.close()
    for k, c in enumerate((9, 50)):
        o = oracle_for(oracle, one)
        o.set_params(gamma=np.full(one.n, gam[c, big]), cfa=np.full(one.n, cfa[c, big]))
        f = cand_forcing(bs, big)
        for y in range(3):
            o.run(np.ascontiguousarray(f[365 * y:365 * (y + 1)]), 1)
            assert np.isclose(q1[k, 0, y], o.station_discharge()[0], rtol=1e-10, atol=1e-14), (k, y)
            o.reset_station_discharge()
        assert_close(st[k], o.get_state(), label=f"largest basin candidate {k}:")


def make_one(bs, b):
    from watergap2_amd.calibration import BasinSet
    return BasinSet([bs.worlds[b]])


def cand_forcing(bs, b):
    """The forcing BasinRunner gives basin b of the set (seed 11, three years)."""
    return bs.worlds[b].forcing(seed=11 + 7919 * b, day0=0, n_days=365 * 3)
