"""Ten years of the global grid (67,420 cells, 3,650 days, the library's two
partitions) on the GPU against the oracle on every cell and station.  About ten
minutes of CPU oracle time on 16 threads, so it runs only with WG_LONG=1 in the
environment; the default suite checks ten years on whole basins cut from the grid."""
import os

import numpy as np
import pytest

from parity import assert_close

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("WG_LONG") != "1", reason="long run: set WG_LONG=1")]


def test_global_grid_ten_years_every_cell(wg, oracle, global_world):
    w = global_world
    years = 10
    print(f"case: global grid {w.n} cells, {years} years, every cell against the oracle", flush=True)
    f = w.forcing(seed=4, day0=0, n_days=365 * years)
    m = wg.WaterBalanceModel(w)
    assert len(m.partitions()) == 2
    m.set_stations(w.basin_outlet)
    m.load_forcing(f)
    m.step_days(0, 1, 365 * years)
    o = oracle.Oracle(w)
    o.set_threads(min(16, os.cpu_count() or 1))
    o.set_stations(w.basin_outlet)
    for y in range(years):
        o.run(np.ascontiguousarray(f[365 * y:365 * (y + 1)]), 1)
        print(f"oracle year {y + 1} done", flush=True)
    st = m.get_state()
    ref = o.get_state()
    assert_close(st, ref, label="year 10:")
    exact = [k for k in ref if k != "started" and np.array_equal(np.asarray(st[k]), np.asarray(ref[k]))]
    print(f"bit-identical fields: {len(exact)} of {len(ref) - 1}", flush=True)
    tot, _ = m.station_discharge()
    assert_close({"discharge": tot[0]}, {"discharge": o.station_discharge()}, label="stations:")
