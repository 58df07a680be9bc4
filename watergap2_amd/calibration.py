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
    path below a taken station are marked as they are reached).  Returns lists of
    0-based cells."""
    count = upstream_cells(dem)
    ds = dem.downstream.astype(np.int64)
    taken = np.zeros(dem.n, dtype=bool)
    sizes = np.exp(rng.uniform(math.log(min_cells), math.log(max_cells), n_basins))
    in_range = (count >= min_cells) & (count <= max_cells)
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

    def run(self, gamma, cfa=None, n_years=30):
        """gamma, cfa: [n_members][n_basins].  Returns annual discharge
        [n_members][n_basins][n_years] (km3/yr, float64)."""
        b = self.basins
        M = self.n_members
        gamma = np.asarray(gamma, dtype=np.float64).reshape(M, b.n_basins)
        g = np.stack([b.per_basin(gamma[m]) for m in range(M)])
        if cfa is None:
            c = np.tile(self.base_cfa, (M, 1))
        else:
            cfa = np.asarray(cfa, dtype=np.float64).reshape(M, b.n_basins)
            c = np.stack([b.per_basin(cfa[m]) for m in range(M)])
        self.model.set_member_params(gamma=g, cfa=c)
        self.model.set_state(self.initial, member=-1)
        self.model.reset_station_discharge()
        out = np.zeros((M, b.n_basins, n_years))
        for y in range(n_years):
            self.model.step_days((y % self.forcing_years) * 365, 1, 365)
            tot, _ = self.model.station_discharge()
            out[:, :, y] = tot
            self.model.reset_station_discharge()
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
    return {"gamma": final, "status": status, "cfa": cfa, "runs": runs_needed, "model_runs": n_runs,
            "unfinished": int(active.sum()), "trajectories": trajectories}


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
