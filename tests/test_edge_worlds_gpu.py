"""Degenerate grids on the GPU against the oracle: a one-cell grid, a two-cell
chain, a 2 x 3 grid, five one-cell basins side by side (no cell has an upstream
cell, so the chain kernels route nothing), and an inland sink basin; with and
without stations, several members; a call of zero days is refused."""
import numpy as np
import pytest

from parity import assert_close

pytestmark = pytest.mark.gpu


def worlds(wg):
    from watergap2_amd.calibration import BasinSet
    out = [(wg.World(wg.regional_spec(seed=1, nrows=1, ncols=1)), "one cell"),
           (wg.World(wg.regional_spec(seed=1, nrows=1, ncols=2)), "two-cell chain"),
           (wg.World(wg.regional_spec(seed=1, nrows=2, ncols=3)), "2 x 3")]
    g = wg.World(wg.global_spec(seed=3, n_land=3000, n_outlets=30, n_sinks=15))
    single = [i for i in range(g.n) if g.downstream[i] <= 0 and (g.inflow9[i] <= 0).all()]
    out.append((BasinSet([g.subset(np.array([c], np.int32)) for c in single[:5]]), "five one-cell basins"))
    sinks = np.flatnonzero(np.asarray(g.ldd) == -1)
    from watergap2_amd.calibration import catchment
    big = max(sinks, key=lambda s: len(catchment(g, s)))
    out.append((g.subset(catchment(g, big)), "inland sink basin"))
    return out


@pytest.mark.parametrize("stations", [True, False])
def test_degenerate_grids_match_oracle(wg, oracle, stations):
    for w, label in worlds(wg):
        print(f"case: {label}, {w.n} cells, stations {stations}, 2 members, 40 days", flush=True)
        n_days, M = 40, 2
        f = w.forcing(seed=5, day0=0, n_days=n_days)
        rng = np.random.default_rng(3)
        gamma = np.tile(w.params[0], (M, 1)) * rng.uniform(0.8, 1.2, (M, w.n))
        cfa = rng.uniform(0.8, 1.2, (M, w.n))
        smax = np.tile(w.smax, (M, 1)).astype(np.float32)
        m = wg.WaterBalanceModel(w, n_members=M)
        m.set_member_params(gamma, cfa, smax)
        if stations:
            m.set_stations(w.basin_outlet)
        m.load_forcing(f)
        with pytest.raises(wg.WaterGapError) as e:   # zero days are refused, nothing runs
            m.step_days(0, 1, 0)
        assert e.value.code == -1
        m.step_days(0, 1, 25)
        m.step_days(25, 26, n_days - 25)
        tot, _ = m.station_discharge()
        for k in range(M):
            o = oracle.Oracle(w)
            o.set_params(gamma[k], cfa[k], smax[k])
            if stations:
                o.set_stations(w.basin_outlet)
            o.run(f, 1)
            assert_close(m.get_state(k), o.get_state(), label=f"{label} m{k}:")
            if stations:
                assert_close({"discharge": tot[k]}, {"discharge": o.station_discharge()}, label=f"{label} m{k}:")
