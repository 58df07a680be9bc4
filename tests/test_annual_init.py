"""routingClass::annualInit for a later year (routing.cpp:979-1495): reservoirs that
start operating in the second year.  The oracle's transfer of the lost land's water
to the reservoir conserves the continental water volume; the library's Parts 1-2
helper selects the reservoirs in operation; on the GPU the second year after
wg_annual_init equals the oracle's."""
import types

import numpy as np
import pytest

from parity import assert_close


def year_worlds(wg, seed=5):
    """The edge-case regional world in its second year, and a first-year copy in
    which a third of the reservoirs (start year 2) do not operate yet."""
    w = wg.World(wg.regional_spec(seed=seed, nrows=30, ncols=40, edge_cases=True, glolake_every=20, res_every=25))
    res = np.flatnonzero(w.reservoir_area > 0)
    late = res[::3]
    start_year = np.where(w.reservoir_area > 0, 1, 0).astype(np.int32)
    start_year[late] = 2

    def copy(year):
        d = {name: getattr(w, name) for name, _ in wg._abi.STATIC_FIELDS if name != "lct"}
        d["lct"] = w.lct
        area, cap = wg.reservoir_year(year, w.reservoir_area, w.stor_cap, start_year)
        d["reservoir_area"], d["stor_cap"] = area, cap
        glo_res = w.glo_res.copy()
        glo_res[start_year > year] = 0.0
        d["glo_res"] = glo_res
        ns = types.SimpleNamespace(n=w.n, **d)
        ns.static_fields = lambda: wg.World.static_fields(ns)
        return ns

    return w, copy(1), copy(2), late


def test_reservoir_year_selects_operating_reservoirs(wg):
    full = np.array([5.0, 3.0, 0.0, 7.0])
    cap = np.array([1.0, 2.0, 0.0, 4.0])
    start = np.array([1990, 2005, 0, 2010], np.int32)
    reg = np.array([0, 0, 0, 1], np.int8)
    area, c = wg.reservoir_year(2000, full, cap, start, reg)
    assert area.tolist() == [5.0, 0.0, 0.0, 7.0] and c.tolist() == [1.0, 0.0, 0.0, 4.0]
    area, c = wg.reservoir_year(2005, full, cap, start)
    assert area.tolist() == [5.0, 3.0, 0.0, 0.0]


def test_oracle_transfer_known_answers(wg, oracle):
    """Part 4 traced by hand from routing.cpp:1313-1421 on chosen states: cell a's
    reservoir fraction grows from 0 to 5 %, cell b's from 3 % to 5 %, both assigned
    to cell a's reservoir.  The land fraction falls by the new fraction (a) or by the
    growth (b); the reservoir gains (canopy + soil + snow) mm x area x (fraction lost
    against the day before) / 100 / 1e6 km3 from each; untouched cells stay.  A cell
    whose land fraction ends above the day before's is refused."""
    w, y1, y2, late = year_worlds(wg)
    a, b = int(late[0]), int(late[1])
    y1.glo_res = y1.glo_res.copy()
    y1.glo_res[b] = 3.0
    y2.glo_res = y1.glo_res.copy()  # only a and b grow
    y2.glo_res[a], y2.glo_res[b] = 5.0, 5.0
    o = oracle.Oracle(y1)
    st = o.get_state()
    for k, v in (("canopy", 2.0), ("soil", 100.0), ("snow", 10.0)):
        st[k][[a, b]] = v
    st["laf"][a], st["laf_prev"][a] = 80.0, 81.0
    st["laf"][b], st["laf_prev"][b] = 60.0, 59.5
    st["reservoir"][a] = 0.25
    o.set_state(st)
    assign = np.arange(1, w.n + 1, dtype=np.int32)
    assign[b] = a + 1
    o.annual_init(y2, assign)
    after = o.get_state()
    assert after["laf"][a] == 75.0 and after["laf"][b] == 58.0
    va = 2.0 * w.cell_area[a] / 1e6 * (6.0 / 100.0) + 100.0 * w.cell_area[a] / 1e6 * (6.0 / 100.0) \
        + 10.0 * w.cell_area[a] / 1e6 * (6.0 / 100.0)
    vb = 2.0 * w.cell_area[b] / 1e6 * (1.5 / 100.0) + 100.0 * w.cell_area[b] / 1e6 * (1.5 / 100.0) \
        + 10.0 * w.cell_area[b] / 1e6 * (1.5 / 100.0)
    assert after["reservoir"][a] == (0.25 + va) + vb
    others = np.ones(w.n, bool)
    others[[a, b]] = False
    for k in ("laf", "reservoir", "soil", "canopy"):
        assert np.array_equal(after[k][others], st[k][others]), k
    # land fraction above the day before after the change: refused
    o2 = oracle.Oracle(y1)
    st2 = o2.get_state()
    st2["laf"][a], st2["laf_prev"][a] = 80.0, 60.0
    o2.set_state(st2)
    with pytest.raises(RuntimeError):
        o2.annual_init(y2, assign)


@pytest.mark.gpu
def test_second_year_with_new_reservoirs_matches_oracle(wg, oracle):
    w, y1, y2, late = year_worlds(wg)
    print(f"case: edge-case regional world, {len(late)} reservoirs starting in year 2, 2 x 365 days", flush=True)
    assign = np.arange(1, w.n + 1, dtype=np.int32)
    f1 = w.forcing(seed=3, day0=0, n_days=365)
    f2 = w.forcing(seed=3, day0=365, n_days=365)
    m = wg.WaterBalanceModel(y1)
    m.set_stations(w.basin_outlet)
    o = oracle.Oracle(y1)
    o.set_stations(w.basin_outlet)
    m.load_forcing(f1)
    m.step_days(0, 1, 365)
    o.run(f1, 1)
    assert_close(m.get_state(), o.get_state(), label="year 1:")
    m.annual_init(y2, assign)
    o.annual_init(y2, assign)
    assert_close(m.get_state(), o.get_state(), label="annual init:")
    m.load_forcing(f2)
    m.step_days(0, 1, 365)
    o.run(f2, 1)
    assert_close(m.get_state(), o.get_state(), label="year 2:")
    tot, _ = m.station_discharge()
    assert_close({"discharge": tot[0]}, {"discharge": o.station_discharge()}, label="station:")
    m.close()
