"""Routing preprocessing (rout_prepare.cpp:61-336): the library's host code against
the oracle's restatement of the reference's sweeps, bit for bit, and the oracle
pinned by brute force on random forests, closed forms and hand-derived known
answers."""
import math

import numpy as np
import pytest

from oracle import wg_oracle
from watergap2_amd import prepare


@pytest.fixture(scope="module")
def gworld(wg):
    w = wg.World(wg.global_spec(seed=1))
    w.ldd_sinks = prepare.replace_flow_to_ocean(w.row, w.col, w.cell_at, w.ldd)
    return w


def random_forest(rng, n, n_roots, loop=False):
    """Random flow graph of n cells: each non-root cell drains to a lower-numbered
    cell in a random order; optionally a 3-cell loop.  Returns (outflow 1-based,
    inflow9, ldd) with LDD 5 at the roots and 1 elsewhere."""
    perm = rng.permutation(n)
    outflow = np.zeros(n, np.int32)
    for k in range(n_roots, n):
        outflow[perm[k]] = perm[rng.integers(0, k)] + 1
    ldd = np.ones(n, np.int8)
    ldd[perm[:n_roots]] = 5
    if loop:
        a, b, c = perm[n - 3:]
        outflow[a], outflow[b], outflow[c] = b + 1, c + 1, a + 1
    inflow9 = np.zeros((n, 9), np.int32)
    slots = np.zeros(n, np.int32)
    for i in range(n):
        d = outflow[i]
        if d > 0:
            # at most 8 inflows by D8; a random graph may exceed it, keep the first 9
            if slots[d - 1] < 9:
                inflow9[d - 1, slots[d - 1]] = i + 1
                slots[d - 1] += 1
    return outflow, inflow9, ldd


def brute_flow_acc(outflow, inflow9):
    """1 + cells upstream, following the listed inflow cells; 0 where a loop lies
    upstream or on the cell."""
    n = len(outflow)
    acc = np.zeros(n, np.int64)
    for i in range(n):
        stack, seen, ok, count = [i], set(), True, 0
        while stack:
            c = stack.pop()
            if c in seen:
                ok = False
                break
            seen.add(c)
            count += 1
            stack.extend(int(u) - 1 for u in inflow9[c] if u > 0)
        acc[i] = count if ok else 0
    return acc


def test_flow_accumulation_brute_force(wg):
    rng = np.random.default_rng(1)
    for trial in range(30):
        outflow, inflow9, _ = random_forest(rng, 200, 5, loop=(trial % 3 == 0))
        expect = brute_flow_acc(outflow, inflow9)
        got = wg_oracle.flow_accumulation(inflow9)
        assert np.array_equal(got, expect.astype(np.int16)), trial
        assert np.array_equal(prepare.flow_accumulation(inflow9), got)


def test_basins_brute_force(wg):
    """Every cell is in the basin of the endpoint it drains to (endpoints numbered
    1.. in cell order), at the number of steps to it; one-cell basins get 0 and the
    others are renumbered in order."""
    rng = np.random.default_rng(2)
    for trial in range(20):
        n = 150
        outflow, inflow9, ldd = random_forest(rng, n, 12)
        ends = np.nonzero(ldd == 5)[0]
        number = {int(e): k + 1 for k, e in enumerate(ends)}
        eb = np.zeros(n, np.int64)
        ed = np.zeros(n, np.int64)
        for i in range(n):
            c, steps = i, 0
            while outflow[c] > 0:
                c, steps = outflow[c] - 1, steps + 1
            eb[i], ed[i] = number[c], steps
        ldd_o, ws, dist, nb, ws2, nm = wg_oracle.derive_basins(ldd, inflow9)
        assert nb == len(ends)
        assert np.array_equal(ws, eb) and np.array_equal(dist, ed)
        sizes = np.bincount(eb, minlength=nb + 1)
        multi = [b for b in range(1, nb + 1) if sizes[b] != 1]
        renum = {b: k + 1 for k, b in enumerate(multi)}
        expect2 = np.array([renum.get(int(b), 0) for b in eb])
        assert nm == len(multi) and np.array_equal(ws2, expect2)
        got = prepare.derive_basins(ldd, inflow9)
        assert all(np.array_equal(a, b) for a, b in zip(got[:3] + got[4:5], (ldd_o, ws, dist, ws2)))
        assert got[3] == nb and got[5] == nm


def test_cell_distances_closed_form(wg):
    """GCELLDIST (rout_prepare.cpp:694-752) on a sphere of radius 6371.211 km:
    north-south 2 pi R 0.5/360, east-west that times sin of the row's colatitude,
    the diagonals sqrt(h_r h_r' + v^2) with the next row north or south."""
    d = wg_oracle.cell_distances().reshape(9, 360).astype(np.float64)
    v = 2 * math.pi * 6371.211 * 0.5 / 360.0
    colat = np.radians(0.25 + 0.5 * np.arange(360))
    h = v * np.sin(colat)
    np.testing.assert_allclose(d[1], v, rtol=1e-6)
    np.testing.assert_allclose(d[7], v, rtol=1e-6)
    # the reference forms the angle in float (l * pi), an absolute error of up to
    # ~6e-7 rad, which near the poles is a large part of sin(l): atol 5e-5 km
    np.testing.assert_allclose(d[3], h, rtol=1e-6, atol=5e-5)
    np.testing.assert_allclose(d[5], h, rtol=1e-6, atol=5e-5)
    assert np.all(d[4] == 0)
    np.testing.assert_allclose(d[0, :359], np.sqrt(h[:359] * h[1:] + v * v), rtol=1e-6, atol=5e-5)
    np.testing.assert_allclose(d[6, 1:], np.sqrt(h[1:] * h[:359] + v * v), rtol=1e-6, atol=5e-5)
    assert d[0, 359] == -99 and d[6, 0] == -99
    assert np.array_equal(d[2], d[0]) and np.array_equal(d[8], d[6])
    assert np.array_equal(prepare.cell_distances(), wg_oracle.cell_distances())


def test_slope_and_length_closed_form(wg, gworld):
    """Slope = altitude drop / (1000 x distance along the LDD x mean meandering
    ratio), at least 1e-5; length = distance to the outflow cell by direction times
    the mean meandering ratio, 55 km x ratio at outlets (rout_prepare.cpp:754-832)."""
    w = gworld
    ldd = w.ldd_sinks
    dist = wg_oracle.cell_distances()
    d9 = dist.reshape(9, 360).astype(np.float64)
    slope = wg_oracle.river_slope(w.meander, dist, w.altitude, w.downstream, ldd, w.row)
    length = wg_oracle.river_length(w.meander, w.row, w.col, dist, w.downstream)
    o = w.downstream.astype(np.int64) - 1
    has = (o >= 0) & (ldd != 5) & (ldd != -1)
    m = w.meander.astype(np.float64)
    r = w.row.astype(np.int64) - 1
    mm = np.where(has, (m + m[np.maximum(o, 0)]) / 2, m)
    drop = w.altitude.astype(np.float64) - w.altitude[np.maximum(o, 0)]
    den = np.where(has, 1000.0 * d9[np.clip(ldd - 1, 0, 8), r] * mm, 1.0)
    expect = np.where(has, drop / den, 0.0)
    expect = np.maximum(expect, 1e-5)
    np.testing.assert_allclose(slope, expect, rtol=2e-6, atol=1e-12)
    same_row = w.row == w.row[np.maximum(o, 0)]
    north = w.row > w.row[np.maximum(o, 0)]
    same_col = w.col == w.col[np.maximum(o, 0)]
    base = np.where(o < 0, 55.0,
                    np.where(same_row, d9[3, r],
                             np.where(north, np.where(same_col, d9[7, r], d9[8, r]),
                                      np.where(same_col, d9[1, r], d9[2, r]))))
    np.testing.assert_allclose(length, base * np.where(o < 0, m, mm), rtol=2e-6)
    # and the library gives the oracle's bits
    assert np.array_equal(prepare.river_slope(w.meander, dist, w.altitude, w.downstream, ldd, w.row), slope)
    assert np.array_equal(prepare.river_length(w.meander, w.row, w.col, dist, w.downstream), length)


def synthetic_reservoirs(w, every=37, seed=5):
    rng = np.random.default_rng(seed)
    area = np.zeros(w.n)
    area[::every] = rng.uniform(5.0, 200.0, len(area[::every]))
    q = rng.uniform(0.1, 50.0, w.n)
    monthly = q[:, None] * (1.0 + 0.8 * np.sin(2 * np.pi * (np.arange(12)[None, :] - rng.integers(0, 12, (w.n, 1))) / 12))
    return area, q, monthly


def test_alloc_coeff_partition_of_unity(wg, gworld):
    """Every non-reservoir cell that reservoirs upstream (within 5 cells, up to the
    next reservoir) serve has its demand split among them: the coefficients of one
    cell sum to 1, and each is the reservoir's mean outflow over their sum."""
    w = gworld
    area, q, _ = synthetic_reservoirs(w)
    ac = wg_oracle.alloc_coeff(w.downstream, area, q)
    served = {}
    for i in np.nonzero(area > 0)[0]:
        d, k = w.downstream[i], 0
        while k < 5 and d > 0 and area[d - 1] <= 0:
            served.setdefault(d - 1, []).append((i, k))
            d, k = w.downstream[d - 1], k + 1
    assert len(served) > 100 and any(len(v) > 1 for v in served.values())
    for cell, lst in served.items():
        total = sum(q[i] for i, _ in lst)
        np.testing.assert_allclose(sum(ac[i, k] for i, k in lst), 1.0, rtol=1e-14)
        for i, k in lst:
            np.testing.assert_allclose(ac[i, k], q[i] / total, rtol=1e-15)
    untouched = np.ones_like(ac, dtype=bool)
    for lst in served.values():
        for i, k in lst:
            untouched[i, k] = False
    assert np.all(ac[untouched] == 1.0)
    assert np.array_equal(prepare.alloc_coeff(w.downstream, area, q), ac)


def test_reservoir_start_month_known_answers(wg):
    """Hand-derived starts of the operational year (rout_prepare.cpp:950-1019): the
    first month of the longest dry period (monthly mean below the annual one),
    scanned from the first wet month; a run over the turn of the year stays whole;
    and the reference's counter, reset only when a run beats the longest so far,
    lets two short runs add up."""
    wet, dry = 2.0, 0.5

    def pattern(dry_months):
        return [dry if m + 1 in dry_months else wet for m in range(12)]

    cases = [
        (pattern({6, 7, 8}), 6),                    # one run, June-August
        (pattern({11, 12, 1, 2}), 11),              # over the turn of the year
        (pattern({3, 4, 7, 8, 9}), 7),              # the longer of two runs
        (pattern({3, 4, 7, 8}), 3),                 # a tie keeps the earlier run
        (pattern({3, 4, 7, 8, 11}), 11),            # 2, then 2 + 1 = 3 (the counter carries)
        (pattern(set(range(1, 13))), 1),            # all dry: January
        (pattern(set()), 1),                        # none dry
    ]
    monthly = np.array([c[0] for c in cases])
    annual = np.ones(len(cases))
    got = wg_oracle.reservoir_start_month(monthly, annual)
    assert list(got) == [c[1] for c in cases]
    assert np.array_equal(prepare.reservoir_start_month(monthly, annual), got)


def test_flowdir_codes(wg):
    codes = np.array([1, 2, 4, 8, 16, 32, 64, 128, 0, -1, 7, 255], np.int16)
    assert list(prepare.ldd_from_flowdir(codes)) == [6, 3, 2, 1, 4, 7, 8, 9, 5, -1, 5, 5]


def test_library_matches_oracle_on_the_global_grid(wg, gworld):
    w = gworld
    ldd = w.ldd_sinks
    acc_o = wg_oracle.flow_accumulation(w.inflow9)
    assert np.array_equal(prepare.flow_accumulation(w.inflow9), acc_o)
    ob = wg_oracle.derive_basins(ldd, w.inflow9)
    lb = prepare.derive_basins(ldd, w.inflow9)
    for a, b in zip(ob, lb):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    area, q, monthly = synthetic_reservoirs(w)
    assert np.array_equal(prepare.reservoir_start_month(monthly, q), wg_oracle.reservoir_start_month(monthly, q))


def test_pipeline_reproduces_the_routing_inputs(wg, gworld):
    """prepare_routing on the grid's flow directions gives the outflow cells,
    inflow cells and routing order that the grid's own generator derived, the flow
    accumulation of the brute force, and one basin per endpoint."""
    w = gworld
    area, q, monthly = synthetic_reservoirs(w)
    out = prepare.prepare_routing(w.row, w.col, w.cell_at, w.altitude, w.meander, ldd=w.ldd_sinks,
                                  reservoir_area=area, mean_outflow=q, mean_outflow_monthly=monthly)
    assert np.array_equal(out["outflow"], w.downstream)
    assert np.array_equal(out["inflow9"], w.inflow9)
    assert np.array_equal(out["rout_order"], w.rout_order)
    assert int(out["flow_acc"].max()) == int(np.bincount(w.basin).max())
    assert int(out["cells_to_outlet"].max()) == w.max_level
    assert out["n_basins"] == int((np.bincount(w.basin) > 1).sum())
    assert np.array_equal(out["alloc_coeff"], wg_oracle.alloc_coeff(w.downstream, area, q))
    assert np.array_equal(out["start_month"], wg_oracle.reservoir_start_month(monthly, q))
    # Arc/Info codes give the same files
    arc = {1: 8, 2: 4, 3: 2, 4: 16, 5: 0, 6: 1, 7: 32, 8: 64, 9: 128, -1: -1}
    flowdir = np.array([arc[int(c)] for c in w.ldd_sinks], np.int16)
    out2 = prepare.prepare_routing(w.row, w.col, w.cell_at, w.altitude, w.meander, flowdir=flowdir)
    for k in ("ldd", "inflow9", "outflow", "flow_acc", "basins", "rout_order", "river_slope", "river_length"):
        assert np.array_equal(out2[k], out[k]), k
