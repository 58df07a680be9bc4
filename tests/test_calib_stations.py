"""Calibration stations on a routing network: the calibration basins of a station
list (cbasinClass::prepare with upstreamBasin, calib_basins.cpp:25-378), the
stations upstream of each basin (upstream_stations.cpp:66-140) and the station
correction factor (calibration.cpp:656-728).  Known answers traced by hand from
the reference's loops, then the library against the oracle's sweep-by-sweep
restatement on random networks and station lists."""
import numpy as np
import pytest

from oracle import calib_oracle as co

# cells 1..6 (1-based): 1 -> 2 -> 3 -> 4 (outlet), 6 -> 5 -> 3
OUTFLOW = [2, 3, 4, 0, 3, 5]
INFLOW = [[0] * 9 for _ in range(6)]
INFLOW[1][0] = 1
INFLOW[2][0], INFLOW[2][1] = 2, 5
INFLOW[3][0] = 3
INFLOW[4][0] = 6


def test_calib_basins_known_answer(wg):
    from watergap2_amd.calibration import calib_basins
    # station 3 claims everything above it; 5 lies in basin 1 and splits it; 4 sits
    # below basin 1; 2 splits basin 1 again; the repeat of 2 and cell 1 (outside the
    # land mask by the reference's test cn <= 1) are skipped
    grid, used, overlap, nb = calib_basins(INFLOW, OUTFLOW, [3, 5, 4, 2, 2, 1])
    assert nb == 4
    assert grid.tolist() == [4, 4, 1, 3, 2, 2]
    assert used.tolist() == [1, 1, 1, 1, 0, 0]
    assert overlap.tolist() == [0, 1, 2, 1, -1, -1]
    # a station whose outflow cell lies in no basin marks it -1
    grid, used, overlap, nb = calib_basins(INFLOW, OUTFLOW, [5])
    assert grid.tolist() == [0, 0, -1, 0, 1, 1] and nb == 1


def test_upstream_stations_known_answer(wg):
    from watergap2_amd.calibration import upstream_stations
    cells = [3, 5, 4, 2]  # basins 1..4
    assert upstream_stations(INFLOW, cells, all_stations=False) == [[2, 4], [], [1], []]
    assert upstream_stations(INFLOW, cells, all_stations=True) == [[2, 4], [], [2, 4, 1], []]


def random_forest(rng, n):
    outflow = np.zeros(n, np.int32)
    for c in range(1, n):
        if rng.random() < 0.9:
            outflow[c] = rng.integers(0, c) + 1  # downstream has a lower index: a forest
    # shuffle the cell numbering so that downstream cells are not always earlier
    perm = rng.permutation(n)
    inv = np.argsort(perm)
    out2 = np.zeros(n, np.int32)
    for c in range(n):
        if outflow[c] > 0:
            out2[inv[c]] = inv[outflow[c] - 1] + 1
    inflow = np.zeros((n, 9), np.int32)
    fill = np.zeros(n, np.int32)
    for c in rng.permutation(n):
        d = out2[c]
        if d > 0 and fill[d - 1] < 9:
            inflow[d - 1, fill[d - 1]] = c + 1
            fill[d - 1] += 1
        elif d > 0:
            out2[c] = 0
    return inflow, out2


def test_library_matches_oracle_on_random_networks(wg):
    from watergap2_amd.calibration import calib_basins, upstream_stations
    rng = np.random.default_rng(12)
    for case in range(40):
        n = int(rng.integers(5, 120))
        inflow, outflow = random_forest(rng, n)
        k = int(rng.integers(1, 15))
        stations = rng.integers(0, n + 1, k).tolist()  # repeats and cells <= 1 included
        grid, used, overlap, nb = calib_basins(inflow, outflow, stations)
        g2, u2, o2 = co.calib_basins(inflow.tolist(), outflow.tolist(), stations)
        assert grid.tolist() == g2.tolist(), case
        assert used.tolist() == u2 and overlap.tolist() == o2, case
        cells = [c for c, u in zip(stations, used) if u]
        for opt in (False, True):
            assert upstream_stations(inflow, cells, opt) == co.upstream_stations(inflow.tolist(), cells, int(opt)), \
                (case, opt)


def test_station_correction_factor(wg):
    from watergap2_amd import _abi
    from watergap2_amd.calibration import station_corr_factor
    lib = _abi.load_library()
    import ctypes as C

    def fresh(ind):
        s = _abi.GammaSearch()
        lib.wg_gamma_search_init(C.byref(s))
        s.cell_corr_factor_ind = ind
        st = co.GammaSearchState()
        st.cell_corr_factor_ind = ind
        return s, st

    meas = np.array([10.0, -99.0, 12.0, 11.0], np.float32)
    sim = np.array([8.0, 5.0, 9.0, 9.5], np.float32)
    # no cell correction: CFS 1, status unchanged
    s, st = fresh(1)
    assert station_corr_factor(s, 2.0, meas, sim) == 1.0 and s.calib_status == 0
    # with it, gamma at the upper limit: CFS = 1.1 * 33 / 26.5 in float
    s, st = fresh(99)
    cfs = station_corr_factor(s, 5.0, meas, sim)
    assert cfs == float(np.float32(float(np.float32(33.0)) * 1.1) / np.float32(26.5))
    assert s.calib_status == 4
    # within 1 % of 1: 1
    s, st = fresh(99)
    assert station_corr_factor(s, 2.0, meas, meas * np.float32(1.005)) == 1.0 and s.calib_status == 0
    rng = np.random.default_rng(3)
    for _ in range(500):
        ind = int(rng.choice([1, 99]))
        s, st = fresh(ind)
        g = float(rng.choice([0.1, 5.0, rng.uniform(0.1, 5.0)]))
        m = rng.uniform(0.1, 50, 8).astype(np.float32)
        m[rng.random(8) < 0.2] = -99.0
        if (m > -1).sum() == 0:
            continue
        sm = (m * rng.uniform(0.8, 1.25, 8)).astype(np.float32)
        a = station_corr_factor(s, g, m, sm)
        b = co.station_corr_factor(st, g, m, sm)
        assert a == float(b) and s.calib_status == st.calib_status


def test_station_edge_cases(wg):
    """An empty station list gives no basin and an all-zero grid; a station list of
    only skipped cells likewise; a basin list with no station upstream gives empty
    lists; bad arguments are refused (WG_E_ARG)."""
    from watergap2_amd.calibration import calib_basins, upstream_stations
    from watergap2_amd import _abi
    grid, used, overlap, nb = calib_basins(INFLOW, OUTFLOW, [])
    assert nb == 0 and not grid.any()
    grid, used, overlap, nb = calib_basins(INFLOW, OUTFLOW, [1, 0, -3])
    assert nb == 0 and not grid.any() and used.tolist() == [0, 0, 0] and overlap.tolist() == [-1, -1, -1]
    for all_st in (0, 1):
        assert upstream_stations(INFLOW, [4], all_st) == [[]]
        assert upstream_stations(INFLOW, [], all_st) == []
    lib = _abi.load_library()
    inflow = np.ascontiguousarray(INFLOW, dtype=np.int32)
    outflow = np.ascontiguousarray(OUTFLOW, dtype=np.int32)
    grid = np.zeros(6, np.int16)
    P32, P16 = _abi.P_i32, _abi.P_i16
    assert lib.wg_calib_basins(0, _abi.ptr(inflow, P32), _abi.ptr(outflow, P32), 0, None, None,
                               _abi.ptr(grid, P16), None) == -1
    assert lib.wg_calib_basins(6, _abi.ptr(inflow, P32), _abi.ptr(outflow, P32), 2, None, None,
                               _abi.ptr(grid, P16), None) == -1
    assert lib.wg_calib_basins(6, _abi.ptr(inflow, P32), _abi.ptr(outflow, P32), 40000,
                               _abi.ptr(np.zeros(40000, np.int32), P32), None, _abi.ptr(grid, P16), None) == -1
    # more list entries than room for them (station 3 lies above station 4)
    lst = np.zeros(1, np.int32)
    cells = np.array([4, 3], np.int32)
    rc = lib.wg_upstream_stations(6, _abi.ptr(inflow, P32), 2, _abi.ptr(cells, P32), 1,
                                  _abi.ptr(np.zeros(3, np.int32), P32), _abi.ptr(lst, P32), 0)
    assert rc == -1
