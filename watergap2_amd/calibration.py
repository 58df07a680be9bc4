"""Batched calibration of the runoff coefficient gamma over many basins.

The reference calibrates one basin per model run: it runs the whole period,
compares the annual discharge at the basin's station with the observations and
asks calibGammaClass::findNewGamma (calibration.cpp:266-636) for the next gamma
(integrateWGHM.cpp:1092-1121), until the 1 % criterion holds or a limit is reached.

Here every basin of a :class:`BasinSet` is one part of a single grid on the GPU:

* :func:`plain_search` runs the reference's search for all basins side by side;
  each iteration is one model run of the whole set with each basin's own gamma
  (basins are independent, so their searches do not interact), and the search
  step itself is the library's ``wg_find_new_gamma``;
* :func:`run_candidates` evaluates K (gamma, CFA) candidates per basin in one run,
  candidates as model members.

Basins are synthetic: disjoint station catchments cut from a seeded global
fractal DEM, sized log-uniformly.  The forcing of a basin is a cycle
of ``forcing_years`` synthetic years repeated over the run (the work per year is
unchanged; forcing memory stays at a few years).
"""
import ctypes as C
import math

import numpy as np

from . import _abi
from . import World, WaterBalanceModel, global_spec, regional_spec, PARAM_INDEX


class BasinSet:
    """Several single-outlet basins concatenated into one grid (reference cell
    order = basin after basin).  Duck-types :class:`World` for the model."""

    def __init__(self, worlds):
        self.worlds = list(worlds)
        self.n_basins = len(self.worlds)
        sizes = [w.n for w in self.worlds]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.n = int(self.offsets[-1])
        self.lct = self.worlds[0].lct
        for name, _ct, _dt, ch in _abi.WORLD_ARRAYS:
            if ch is None or name in ("inflow9", "basin", "level"):
                continue
            parts = [getattr(w, name) for w in self.worlds]
            if ch == "params":
                arr = np.concatenate(parts, axis=1)
            else:
                arr = np.concatenate(parts, axis=0)
            setattr(self, name, np.ascontiguousarray(arr))
        # downstream and routing order are 1-based cell numbers: shift per basin
        ds = []
        ro = []
        outlets = []
        for b, w in enumerate(self.worlds):
            off = int(self.offsets[b])
            d = w.downstream.astype(np.int32).copy()
            d[d > 0] += off
            ds.append(d)
            ro.append(w.rout_order.astype(np.int32) + off)
            assert w.n_basins == 1
            outlets.append(int(w.basin_outlet[0]) + off)
        self.downstream = np.ascontiguousarray(np.concatenate(ds))
        self.rout_order = np.ascontiguousarray(np.concatenate(ro))
        self.basin_outlet = np.array(outlets, dtype=np.int32)
        self.basin = np.repeat(np.arange(self.n_basins, dtype=np.int32), sizes)
        self.max_level = max(w.max_level for w in self.worlds)
        self.mean_nus = None
        self.alloc_coeff = None

    def static_fields(self):
        return World.static_fields(self)

    def forcing(self, seed, day0, n_days):
        """[n_days, 4, N] forcing: each basin's own synthetic series."""
        out = np.empty((n_days, 4, self.n), dtype=np.float32)
        for b, w in enumerate(self.worlds):
            a, e = int(self.offsets[b]), int(self.offsets[b + 1])
            out[:, :, a:e] = w.forcing(seed=seed + 7919 * b, day0=day0, n_days=n_days)
        return out

    def per_basin(self, values):
        """Expands one value per basin to one value per cell."""
        return np.repeat(np.asarray(values, dtype=np.float64), np.diff(self.offsets))


def upstream_cells(world):
    """Number of cells draining through each cell, itself included."""
    order = np.argsort(world.rout_order)  # upstream cells come first
    count = np.ones(world.n, dtype=np.int64)
    ds = world.downstream.astype(np.int64)
    for i in order:
        if ds[i] > 0:
            count[ds[i] - 1] += count[i]
    return count


def catchment(world, station):
    """0-based cells upstream of (and including) a station cell, ascending."""
    inflow = world.inflow9
    stack, out = [int(station)], []
    while stack:
        c = stack.pop()
        out.append(c)
        stack.extend(int(u) - 1 for u in inflow[c] if u > 0)
    return np.sort(np.array(out, dtype=np.int32))


def cut_basins(dem, n_basins, rng, min_cells, max_cells):
    """Up to n_basins disjoint catchments of one grid with cell counts log-uniform
    in [min_cells, max_cells]: for each drawn size (largest first) the station is
    the free cell whose catchment count is nearest.  A cell is free when it lies in
    no basin taken before and no such basin lies upstream of it (cells on the flow
    path below a taken station are marked as they are reached).  Inland sinks are
    never stations: their discharge is not recorded (routing.cpp:4232-4238), so
    such a basin has nothing to calibrate against.  Returns lists of 0-based cells."""
    count = upstream_cells(dem)
    ds = dem.downstream.astype(np.int64)
    taken = np.zeros(dem.n, dtype=bool)
    sizes = np.exp(rng.uniform(math.log(min_cells), math.log(max_cells), n_basins))
    in_range = (count >= min_cells) & (count <= max_cells) & (np.asarray(dem.ldd) != -1)
    out = []
    for size in sorted(sizes, reverse=True):
        cand = np.flatnonzero(in_range & ~taken)
        if len(cand) == 0:
            break
        st = int(cand[np.argmin(np.abs(np.log(count[cand] / size)))])
        cells = catchment(dem, st)
        taken[cells] = True
        c = ds[st] - 1
        while c >= 0 and not taken[c]:
            taken[c] = True
            c = ds[c] - 1
        out.append(cells)
    return out


def make_basins(n_basins, seed, min_cells=50, max_cells=5000, dem_cells=200000):
    """Synthetic single-outlet basins with log-uniform cell counts in
    [min_cells, max_cells], cut as disjoint station catchments from a seeded global
    0.5-degree DEM of dem_cells land cells (a further DEM, seed + 1, ..., when one
    has too few).  Basins come in random order."""
    rng = np.random.default_rng(seed)
    worlds = []
    dem_seed = int(seed) * 7 + 3
    while len(worlds) < n_basins:
        dem = World(global_spec(seed=dem_seed, n_land=dem_cells, n_outlets=max(40, dem_cells // 500),
                                n_sinks=max(4, dem_cells // 5000)))
        need = n_basins - len(worlds)
        for cells in cut_basins(dem, need, rng, min_cells, max_cells):
            worlds.append(dem.subset(cells))
        dem_seed += 1
    order = rng.permutation(len(worlds))
    return BasinSet([worlds[i] for i in order])


def make_lattice_basins(n_basins, seed, min_cells=50, max_cells=5000):
    """Synthetic single-outlet basins with log-uniform cell counts in
    [min_cells, max_cells], each on its own DEM seed and latitude band."""
    rng = np.random.default_rng(seed)
    worlds = []
    for b in range(n_basins):
        size = int(round(math.exp(rng.uniform(math.log(min_cells), math.log(max_cells)))))
        ncols = max(2, int(round(math.sqrt(size * rng.uniform(0.6, 1.6)))))
        nrows = max(2, int(round(size / ncols)))
        row0 = int(rng.integers(30, 300 - nrows)) if nrows < 260 else 20
        col0 = int(rng.integers(1, 720 - ncols))
        spec = regional_spec(seed=int(seed * 100003 + b), nrows=nrows, ncols=ncols, row0=row0, col0=col0)
        worlds.append(World(spec))
    return BasinSet(worlds)


class BasinRunner:
    """One model handle over a basin set: resettable runs of n_years years with
    per-member, per-basin gamma and CFA; returns annual station discharge."""

    def __init__(self, basins, n_members=1, forcing_years=3, seed=11, forcing=None):
        self.basins = basins
        self.n_members = n_members
        self.forcing_years = forcing_years
        self.model = WaterBalanceModel(basins, n_members=n_members)
        self.model.set_stations(basins.basin_outlet)
        if forcing is None:
            forcing = basins.forcing(seed=seed, day0=0, n_days=365 * forcing_years)
        assert forcing.shape[0] == 365 * forcing_years
        self.forcing = forcing
        self.model.load_forcing(forcing)
        self.initial = self.model.get_state(0)
        self.base_gamma = basins.params[PARAM_INDEX["P_GAMRUN_C"]].copy()
        self.base_cfa = basins.params[PARAM_INDEX["P_CFA"]].copy()

    def run(self, gamma, cfa=None, n_years=30, cell_cfa=None, pot=False):
        """gamma, cfa: [n_members][n_basins]; cell_cfa: [n_members][N] cell
        correction factors in place of cfa (a correction grid).  Returns annual
        discharge [n_members][n_basins][n_years] (km3/yr, float64); with pot, also
        each year's potential cell runoff [n_members][n_years][N] as the reference
        writes it (G_POT_CELL_RUNOFF_<year>.UNF0, float)."""
        b = self.basins
        M = self.n_members
        gamma = np.asarray(gamma, dtype=np.float64).reshape(M, b.n_basins)
        g = np.stack([b.per_basin(gamma[m]) for m in range(M)])
        if cell_cfa is not None:
            c = np.asarray(cell_cfa, dtype=np.float64).reshape(M, b.n)
        elif cfa is None:
            c = np.tile(self.base_cfa, (M, 1))
        else:
            cfa = np.asarray(cfa, dtype=np.float64).reshape(M, b.n_basins)
            c = np.stack([b.per_basin(cfa[m]) for m in range(M)])
        self.model.set_member_params(gamma=g, cfa=c)
        self.model.set_state(self.initial, member=-1)
        self.model.reset_station_discharge()
        self.model.pot_cell_runoff_output(pot)
        out = np.zeros((M, b.n_basins, n_years))
        pot_years = np.zeros((M, n_years, b.n), dtype=np.float32) if pot else None
        for y in range(n_years):
            self.model.step_days((y % self.forcing_years) * 365, 1, 365)
            tot, _ = self.model.station_discharge()
            out[:, :, y] = tot
            self.model.reset_station_discharge()
            if pot:
                for m in range(M):
                    pot_years[m, y] = self.model.pot_cell_runoff(m).astype(np.float32)
                self.model.pot_cell_runoff_output(True)  # the next year starts at 0 (routing.cpp:1009)
        if pot:
            self.model.pot_cell_runoff_output(False)
            return out, pot_years
        return out

    def close(self):
        self.model.close()


def _search_lib():
    lib = _abi.load_library()
    return lib


def plain_search(runner, measured, n_years=30, max_iter=40, gamma_from_file=None):
    """The reference's gamma search (findNewGamma) for every basin of the
    runner's set, all basins advancing together, one model run per iteration.

    measured: [n_basins][n_years] observed annual discharge (<= -1: no data).
    Returns a dict with the final gamma per basin (the last gamma run, as the
    reference's test run uses), the search status per basin, the number of model
    runs each basin needed and the total number of model runs."""
    assert runner.n_members == 1
    lib = _search_lib()
    nb = runner.basins.n_basins
    measured = np.asarray(measured, dtype=np.float32).reshape(nb, n_years)
    states = [_abi.GammaSearch() for _ in range(nb)]
    gamma = np.zeros(nb, dtype=np.float32)
    for i, s in enumerate(states):
        lib.wg_gamma_search_init(C.byref(s))
        g0 = -1.0 if gamma_from_file is None else float(gamma_from_file[i])
        gamma[i] = lib.wg_gamma_start(C.byref(s), g0)
    active = np.ones(nb, dtype=bool)
    final = gamma.copy()
    runs_needed = np.zeros(nb, dtype=np.int32)
    trajectories = [[float(g)] for g in gamma]
    n_runs = 0
    while active.any() and n_runs < max_iter:
        q = runner.run(gamma.astype(np.float64)[None, :], n_years=n_years)[0]
        n_runs += 1
        for i in np.flatnonzero(active):
            runs_needed[i] += 1
            sim = np.ascontiguousarray(q[i].astype(np.float32))
            meas = np.ascontiguousarray(measured[i])
            g_new = lib.wg_find_new_gamma(C.byref(states[i]), float(gamma[i]), n_years,
                                          _abi.ptr(meas, _abi.P_f), _abi.ptr(sim, _abi.P_f), None, None)
            if g_new < 0:
                active[i] = False
                final[i] = gamma[i]
            else:
                gamma[i] = np.float32(g_new)
                trajectories[i].append(float(g_new))
    status = np.array([s.calib_status for s in states], dtype=np.int32)
    cfa = np.array([s.cell_corr_factor for s in states], dtype=np.float32)
    sim_avg = np.array([s.sim_runoff_avg for s in states], dtype=np.float32)
    meas_adapt = np.array([s.measured_avg_adapt for s in states], dtype=np.float32)
    return {"gamma": final, "status": status, "cfa": cfa, "runs": runs_needed, "model_runs": n_runs,
            "unfinished": int(active.sum()), "trajectories": trajectories,
            "sim_avg": sim_avg, "measured_adapt": meas_adapt}


def correction_grid(calib_basin, station_number, annual_pot, simulated, measured, corr):
    """Cell correction factors of one calibration basin (wg_cell_corr_factor_grid,
    createCorrectionGrid, calibration.cpp:730-801).  annual_pot: [n_years][N] float;
    corr: the current factors [N], returned updated for the basin's cells, with
    the mean potential runoff [N]."""
    lib = _abi.load_library()
    basin = np.ascontiguousarray(calib_basin, dtype=np.int16)
    annual = np.ascontiguousarray(annual_pot, dtype=np.float32)
    out = np.array(corr, dtype=np.float64)
    mean = np.zeros(annual.shape[1], np.float32)
    rc = lib.wg_cell_corr_factor_grid(annual.shape[1], _abi.ptr(basin, _abi.P_i16), int(station_number),
                                      annual.shape[0], _abi.ptr(annual, _abi.P_f), float(simulated), float(measured),
                                      _abi.ptr(mean, _abi.P_f), _abi.ptr(out, _abi.P_d))
    if rc != 0:
        raise RuntimeError(f"wg_cell_corr_factor_grid failed: {_abi.ERRORS.get(rc, rc)}")
    return out, mean


def correction_grids(runner, result, n_years=30):
    """The cell correction grid after a plain search: for the basins whose search
    ended with a cell correction factor (status 3, calibration.cpp:625-628), one
    run of the final gamma keeps the annual potential cell runoff (the files the
    reference's last run wrote), and each such basin's cells get
    createCorrectionGrid's factors.  Returns the per-cell factors [N] (the runner's
    CFA elsewhere) and the basins corrected."""
    b = runner.basins
    corr = runner.base_cfa.astype(np.float64).copy()
    todo = np.flatnonzero(result["status"] == 3)
    if len(todo) == 0:
        return corr, todo
    _, pot = runner.run(np.asarray(result["gamma"], np.float64)[None, :], n_years=n_years, pot=True)
    calib_basin = np.repeat(np.arange(1, b.n_basins + 1, dtype=np.int16), np.diff(b.offsets))
    for i in todo:
        corr, _ = correction_grid(calib_basin, i + 1, pot[0], result["sim_avg"][i], result["measured_adapt"][i],
                                  corr)
    return corr, todo


def within_1pc(sim, measured):
    """The reference's 1 % criterion on annual series (calibration.cpp:327)."""
    sim = np.asarray(sim, dtype=np.float32)
    measured = np.asarray(measured, dtype=np.float32)
    ok = measured > -1
    d = np.float32(0)
    ms = np.float32(0)
    for s_, m_ in zip(sim[ok], measured[ok]):
        d = np.float32(d + np.float32(s_ - m_))
        ms = np.float32(ms + m_)
    ny = int(ok.sum())
    avg = np.float32(ms / np.float32(ny))
    return abs(float(np.float32(d / np.float32(ny * avg)))) < 0.01


def calib_basins(inflow9, outflow, station_cells):
    """Calibration basins of station cells (1-based, station-file order) on a
    routing network (cbasinClass::prepare, calib_basins.cpp:25-300): returns
    (calib_basin [N] int16, used [K], overlap [K], n_basins).  See
    wg_calib_basins in include/watergap2_amd.h."""
    lib = _abi.load_library()
    inflow9 = np.ascontiguousarray(inflow9, dtype=np.int32)
    outflow = np.ascontiguousarray(outflow, dtype=np.int32)
    cells = np.ascontiguousarray(station_cells, dtype=np.int32)
    n = len(outflow)
    grid = np.zeros(n, np.int16)
    used = np.zeros(len(cells), np.int32)
    overlap = np.zeros(len(cells), np.int32)
    rc = lib.wg_calib_basins(n, _abi.ptr(inflow9, _abi.P_i32), _abi.ptr(outflow, _abi.P_i32), len(cells),
                             _abi.ptr(cells, _abi.P_i32), _abi.ptr(used, _abi.P_i32), _abi.ptr(grid, _abi.P_i16),
                             _abi.ptr(overlap, _abi.P_i32))
    if rc < 0:
        raise RuntimeError(f"wg_calib_basins failed: {_abi.ERRORS.get(rc, rc)}")
    return grid, used, overlap, rc


def upstream_stations(inflow9, basin_cells, all_stations=False):
    """Per basin, the basins whose stations lie upstream of its station
    (upstream_stations.cpp:66-140): directly upstream, or all with all_stations.
    Returns a list of lists of 1-based basin numbers in the reference's order."""
    lib = _abi.load_library()
    inflow9 = np.ascontiguousarray(inflow9, dtype=np.int32)
    cells = np.ascontiguousarray(basin_cells, dtype=np.int32)
    nb = len(cells)
    cap = max(1, nb * nb)
    offsets = np.zeros(nb + 1, np.int32)
    out = np.zeros(cap, np.int32)
    rc = lib.wg_upstream_stations(len(inflow9), _abi.ptr(inflow9, _abi.P_i32), nb, _abi.ptr(cells, _abi.P_i32),
                                  1 if all_stations else 0, _abi.ptr(offsets, _abi.P_i32), _abi.ptr(out, _abi.P_i32),
                                  cap)
    if rc != 0:
        raise RuntimeError(f"wg_upstream_stations failed: {_abi.ERRORS.get(rc, rc)}")
    return [out[offsets[b]:offsets[b + 1]].tolist() for b in range(nb)]


def station_corr_factor(search, gamma, measured, simulated):
    """The station correction factor CFS after a search (writeCorrFactors,
    calibration.cpp:656-728); `search` is the basin's _abi.GammaSearch."""
    lib = _abi.load_library()
    meas = np.ascontiguousarray(measured, dtype=np.float32)
    sim = np.ascontiguousarray(simulated, dtype=np.float32)
    return float(lib.wg_station_corr_factor(C.byref(search), float(gamma), len(meas), _abi.ptr(meas, _abi.P_f),
                                            _abi.ptr(sim, _abi.P_f)))


class StationRunner:
    """One model handle over a routing network with calibration stations, which may
    be nested: each station's calibration basin is the part of its catchment below
    the stations upstream of it (calib_basins.cpp:25-300).  gamma and CFA are set per
    calibration basin; cells in no basin keep the grid's own values.  run() returns
    the annual discharge at every station."""

    def __init__(self, world, station_cells, n_members=1, forcing_years=3, seed=11, forcing=None):
        self.world = world
        grid, used, _overlap, nb = calib_basins(world.inflow9, world.downstream, station_cells)
        self.calib_basin = grid
        self.cells = np.array([c for c, u in zip(station_cells, used) if u], dtype=np.int32)  # 1-based
        self.n_basins = nb
        self.direct_upstream = upstream_stations(world.inflow9, self.cells, all_stations=False)
        self.n_members = n_members
        self.forcing_years = forcing_years
        self.model = WaterBalanceModel(world, n_members=n_members)
        self.model.set_stations(self.cells - 1)
        if forcing is None:
            forcing = world.forcing(seed=seed, day0=0, n_days=365 * forcing_years)
        self.forcing = forcing
        self.model.load_forcing(forcing)
        self.initial = self.model.get_state(0)
        self.base_gamma = world.params[PARAM_INDEX["P_GAMRUN_C"]].copy()
        self.base_cfa = world.params[PARAM_INDEX["P_CFA"]].copy()

    def per_cell(self, values, base):
        """Per-basin values (1-based basins) on the cells of each calibration basin."""
        out = base.copy()
        inside = self.calib_basin > 0
        out[inside] = np.asarray(values, dtype=np.float64)[self.calib_basin[inside] - 1]
        return out

    def run(self, gamma, cfa=None, n_years=30):
        """gamma, cfa: [n_members][n_basins].  Returns annual station discharge
        [n_members][n_basins][n_years] (km3/yr, float64: each the sum of the day's
        values in day order, routing.cpp:7232-7239)."""
        M, nb = self.n_members, self.n_basins
        gamma = np.asarray(gamma, dtype=np.float64).reshape(M, nb)
        g = np.stack([self.per_cell(gamma[m], self.base_gamma) for m in range(M)])
        if cfa is None:
            c = np.tile(self.base_cfa, (M, 1))
        else:
            cfa = np.asarray(cfa, dtype=np.float64).reshape(M, nb)
            c = np.stack([self.per_cell(cfa[m], self.base_cfa) for m in range(M)])
        self.model.set_member_params(gamma=g, cfa=c)
        self.model.set_state(self.initial, member=-1)
        self.model.reset_station_discharge()
        out = np.zeros((M, nb, n_years))
        for y in range(n_years):
            self.model.step_days((y % self.forcing_years) * 365, 1, 365)
            tot, _ = self.model.station_discharge()
            out[:, :, y] = tot
            self.model.reset_station_discharge()
        return out

    def upstream_inflow(self, annual):
        """Annual inflow from each basin's direct upstream stations
        (getAnnualUpstreamInflow, routing.cpp:7241-7258): the stations' annual
        discharges added in list order, as float (calibGammaClass::setUpstInflow)."""
        nb, ny = annual.shape
        out = np.zeros((nb, ny), dtype=np.float32)
        for b in range(nb):
            for y in range(ny):
                total = 0.0
                for u in self.direct_upstream[b]:
                    total += annual[u - 1, y]
                out[b, y] = np.float32(total)
        return out

    def close(self):
        self.model.close()


def nested_search(runner, measured, n_years=30, max_iter=40):
    """The reference's calibration of every station of a network, upstream first:
    a station's search starts once the stations directly upstream of it are done,
    and it compares its station's discharge with the observations given the
    simulated inflow from those stations (calibration.cpp:266-636 with
    integrateWGHM.cpp:967-974).  Stations of one tier search side by side, one
    model run of the whole network per step; a finished station keeps its last
    gamma.  After a station's search the test run's discharge gives its station
    correction factor (calibration.cpp:656-728).  Returns the final gamma, status,
    CFA and CFS per station, the runs each needed, the tiers and the number of
    model runs."""
    assert runner.n_members == 1
    lib = _abi.load_library()
    nb = runner.n_basins
    measured = np.asarray(measured, dtype=np.float32).reshape(nb, n_years)
    states = [_abi.GammaSearch() for _ in range(nb)]
    gamma = np.ones(nb, dtype=np.float32)
    for i, s in enumerate(states):
        lib.wg_gamma_search_init(C.byref(s))
        gamma[i] = lib.wg_gamma_start(C.byref(s), -1.0)
    # tier of a station: 0 without upstream stations, else 1 + the deepest tier above
    tier = np.full(nb, -1)
    while (tier < 0).any():
        for b in range(nb):
            ups = runner.direct_upstream[b]
            if tier[b] < 0 and all(tier[u - 1] >= 0 for u in ups):
                tier[b] = 1 + max((tier[u - 1] for u in ups), default=-1)
    status = np.zeros(nb, dtype=np.int32)
    runs = np.zeros(nb, dtype=np.int32)
    cfs = np.ones(nb, dtype=np.float32)
    n_runs = 0
    for t in range(int(tier.max()) + 1):
        active = tier == t
        for _ in range(max_iter):
            if not active.any():
                break
            q = runner.run(gamma.astype(np.float64)[None, :], n_years=n_years)[0]
            n_runs += 1
            inflow = runner.upstream_inflow(q)
            for i in np.flatnonzero(active):
                runs[i] += 1
                sim = np.ascontiguousarray(q[i].astype(np.float32))
                meas = np.ascontiguousarray(measured[i])
                infl = np.ascontiguousarray(inflow[i])
                g_new = lib.wg_find_new_gamma(C.byref(states[i]), float(gamma[i]), n_years, _abi.ptr(meas, _abi.P_f),
                                              _abi.ptr(sim, _abi.P_f), _abi.ptr(infl, _abi.P_f), None)
                if g_new < 0:
                    active[i] = False
                else:
                    gamma[i] = np.float32(g_new)
        # the test runs with the final gammas give the station correction factors
        q = runner.run(gamma.astype(np.float64)[None, :], n_years=n_years)[0]
        n_runs += 1
        for i in np.flatnonzero(tier == t):
            sim = np.ascontiguousarray(q[i].astype(np.float32))
            cfs[i] = station_corr_factor(states[i], gamma[i], measured[i], sim)
    status = np.array([s.calib_status for s in states], dtype=np.int32)
    cfa = np.array([s.cell_corr_factor for s in states], dtype=np.float32)
    return {"gamma": gamma, "status": status, "cfa": cfa, "cfs": cfs, "runs": runs, "tier": tier,
            "model_runs": n_runs}
