"""Batched calibration on the GPU against the same search driven by the CPU oracle."""
import numpy as np
import pytest

from oracle import calib_oracle as co

pytestmark = pytest.mark.gpu


class OracleRunner:
    """The CPU oracle over a basin set, run like watergap2_amd.calibration.BasinRunner."""

    def __init__(self, oracle, basins, forcing_years=3, seed=11):
        self.basins = basins
        self.o = oracle.Oracle(basins)
        self.o.set_stations(basins.basin_outlet)
        self.f = basins.forcing(seed=seed, day0=0, n_days=365 * forcing_years)
        self.forcing_years = forcing_years
        self.initial = self.o.get_state()

    def run(self, gamma, cfa, n_years, cell_cfa=None, pot=False):
        b = self.basins
        from watergap2_amd import PARAM_INDEX
        if cell_cfa is None:
            cell_cfa = b.params[PARAM_INDEX["P_CFA"]] if cfa is None else b.per_basin(cfa)
        self.o.set_params(gamma=b.per_basin(gamma), cfa=cell_cfa)
        self.o.set_state(self.initial)
        self.o.reset_station_discharge()
        self.o.reset_pot_cell_runoff()
        out = np.zeros((b.n_basins, n_years))
        pots = np.zeros((n_years, b.n), np.float32)
        for y in range(n_years):
            k = y % self.forcing_years
            self.o.run(self.f[365 * k:365 * (k + 1)], 1)
            out[:, y] = self.o.station_discharge()
            self.o.reset_station_discharge()
            pots[y] = self.o.pot_cell_runoff().astype(np.float32)
            self.o.reset_pot_cell_runoff()
        return (out, pots) if pot else out


def test_candidates_match_oracle(wg, oracle):
    """K (gamma, CFA) candidates per basin as members of one run: every member's
    annual station discharge equals the oracle run of that candidate."""
    from watergap2_amd.calibration import BasinRunner, make_basins
    print("case: 3 basins x 3 candidates, 3 years", flush=True)
    bs = make_basins(3, seed=21, min_cells=30, max_cells=150)
    K, years = 3, 3
    rng = np.random.default_rng(4)
    gam = rng.uniform(0.1, 5.0, (K, bs.n_basins))
    cfa = rng.uniform(0.5, 1.5, (K, bs.n_basins))
    r = BasinRunner(bs, n_members=K, forcing_years=3)
    q = r.run(gam, cfa, n_years=years)
    ref = OracleRunner(oracle, bs, forcing_years=3)
    for k in range(K):
        qo = ref.run(gam[k], cfa[k], years)
        assert np.array_equal(q[k], qo), k


def test_plain_search_matches_oracle_driven_search(wg, oracle):
    """The batched search on the GPU takes the same steps as the reference's search
    driven by oracle model runs, basin by basin, to the same gamma."""
    from watergap2_amd.calibration import BasinRunner, make_basins, plain_search
    print("case: plain gamma search, 4 basins, 3 years", flush=True)
    bs = make_basins(4, seed=33, min_cells=30, max_cells=150)
    years = 3
    hidden = np.array([0.4, 1.7, 3.1, 4.6])
    ref = OracleRunner(oracle, bs, forcing_years=3)
    cfa1 = None
    measured = ref.run(hidden, cfa1, years).astype(np.float32)
    r = BasinRunner(bs, n_members=1, forcing_years=3)
    res = plain_search(r, measured, n_years=years)
    assert res["unfinished"] == 0
    for b in range(bs.n_basins):
        st = co.GammaSearchState()
        g = co.gamma_start(st, -1.0)
        traj = [float(g)]
        while True:
            sim = ref.run(np.full(bs.n_basins, float(g)), cfa1, years)[b].astype(np.float32)
            g_new = co.find_new_gamma(st, g, measured[b], sim)
            if g_new < 0:
                break
            g = g_new
            traj.append(float(g))
        assert traj == res["trajectories"][b], (b, traj, res["trajectories"][b])
        assert np.float32(res["gamma"][b]) == g


def test_nested_stations_search(wg, oracle):
    """Four nested stations on the 2,000-cell regional grid: the station runner's
    discharges equal the oracle's with the same per-basin gamma, and the tiered
    search (upstream stations first, simulated upstream inflow passed to
    findNewGamma) brings every station within 1 % of its hidden-gamma target."""
    from watergap2_amd.calibration import StationRunner, nested_search, upstream_cells, within_1pc
    w = wg.World(wg.regional_spec(seed=0))
    count = upstream_cells(w)
    outlet = int(np.argmax(count))
    # three interior stations with catchments of 150-700 cells, nested where the network allows
    cand = np.flatnonzero((count > 150) & (count < 700))
    rng = np.random.default_rng(8)
    picks = [int(c) for c in rng.choice(cand, 3, replace=False)]
    stations = [p + 1 for p in picks] + [outlet + 1]
    print(f"case: nested stations {stations} on the regional grid, 3 years", flush=True)
    years = 3
    r = StationRunner(w, stations, forcing_years=3)
    nb = r.n_basins
    assert nb == 4 and (r.calib_basin > 0).sum() == w.n
    hidden = rng.uniform(0.3, 4.0, nb)
    q = r.run(hidden[None, :], n_years=years)[0]
    # the same per-cell parameters in the oracle
    o = oracle.Oracle(w)
    o.set_params(gamma=r.per_cell(hidden, r.base_gamma))
    o.set_stations(r.cells - 1)
    for y in range(years):
        o.run(r.forcing[365 * y:365 * (y + 1)], 1)
        assert np.array_equal(q[:, y], o.station_discharge()), y
        o.reset_station_discharge()
    measured = q.astype(np.float32)
    res = nested_search(r, measured, n_years=years)
    final = r.run(res["gamma"].astype(np.float64)[None, :], n_years=years)[0]
    r.close()
    assert (res["status"] >= 1).all()
    assert all(within_1pc(final[b], measured[b]) for b in range(nb)), res
    assert res["tier"].max() >= 1  # the outlet waits for the stations above it


def test_cell_correction_grid_after_a_search_at_the_limit(wg, oracle):
    """Targets three times the discharge at the lowest gamma: the search stops at
    the lower limit with a cell correction factor (status 3), and the correction
    grid (createCorrectionGrid) is built from the final run's annual potential cell
    runoff.  The GPU's annual potential runoff equals the oracle's, the grid equals
    the float restatement on the oracle's grids, and running with the grid moves
    the discharge towards the targets."""
    from watergap2_amd.calibration import BasinRunner, correction_grids, make_basins, plain_search
    print("case: correction grid, 3 basins, 3 years", flush=True)
    bs = make_basins(3, seed=21, min_cells=30, max_cells=150)
    years = 3
    r = BasinRunner(bs, n_members=1, forcing_years=3)
    low = r.run(np.full((1, bs.n_basins), 0.1), n_years=years)[0]
    measured = (3.0 * low).astype(np.float32)
    res = plain_search(r, measured, n_years=years)
    assert (res["status"] == 3).all(), res["status"]
    assert np.all(res["gamma"] == np.float32(0.1))
    corr, todo = correction_grids(r, res, n_years=years)
    assert list(todo) == [0, 1, 2]
    q, pot = r.run(res["gamma"].astype(np.float64)[None, :], n_years=years, pot=True)
    ref = OracleRunner(oracle, bs, forcing_years=3)
    qo, pot_o = ref.run(res["gamma"].astype(np.float64), None, years, pot=True)
    assert np.array_equal(q[0], qo)
    assert np.array_equal(pot[0], pot_o)
    expect = r.base_cfa.astype(np.float64).copy()
    basin_no = np.repeat(np.arange(1, bs.n_basins + 1), np.diff(bs.offsets))
    for i in range(bs.n_basins):
        expect, _ = co.correction_grid(basin_no, i + 1, pot_o, res["sim_avg"][i], res["measured_adapt"][i],
                                       expect)
    assert np.array_equal(corr, expect)
    assert np.all((corr >= 0.5) & (corr <= 1.5)) and np.any(corr > 1.0)
    q2 = r.run(res["gamma"].astype(np.float64)[None, :], n_years=years, cell_cfa=corr[None, :])[0]
    assert np.all(q2.sum(axis=1) > q[0].sum(axis=1))
    assert np.all(q2.sum(axis=1) < measured.sum(axis=1))
