"""CPU restatement of the calibration search (test infrastructure only).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

find_new_gamma restates calibGammaClass::findNewGamma (calibration.cpp:266-636)
with numpy float32 scalars for the reference's float variables; the function
statics (calibration.cpp:275-284) live in a GammaSearchState object.
"""
import numpy as np

f32 = np.float32


class GammaSearchState:
    """Function statics and members of calibGammaClass used by findNewGamma."""

    def __init__(self):
        self.gamma_low = f32(-99)            # calibration.cpp:282
        self.gamma_high = f32(-99)           # calibration.cpp:283
        self.sum_diff_old = f32(0)           # calibration.cpp:275
        self.call_counter = 0                # calibration.cpp:278
        self.gamma_cond = 0                  # calibration.h:14
        self.calib_status = 0                # calibration.h:15
        self.cell_corr_factor_ind = 1        # calibration.h:13
        self.cell_corr_factor = f32(1)       # calibration.h:48
        self.upper = f32(5.)                 # calibration.cpp:47
        self.lower = f32(0.1)                # calibration.cpp:48
        self.measured_avg_adapt = f32(0)     # measuredRunoffAvgAdapt of the last call
        self.sim_runoff_avg = f32(0)         # simRunoffSum / numberOfYears of the last call


def gamma_start(state, gamma_from_file):
    """calibration.cpp:196-205."""
    g = f32(gamma_from_file)
    if g < 0:
        return f32(1)
    if g > state.upper or g < state.lower:
        return f32(1)
    return g


def find_new_gamma(st, gamma_old, measured, simulated, inflow=None, water_use=None):
    """calibration.cpp:266-636; returns the next gamma (float32) or -99."""
    gamma_old = f32(gamma_old)
    n = len(measured)
    inflow = np.zeros(n, np.float32) if inflow is None else np.asarray(inflow, np.float32)
    water_use = np.zeros(n, np.float32) if water_use is None else np.asarray(water_use, np.float32)
    meas = np.asarray(measured, np.float32)
    sim = np.asarray(simulated, np.float32)
    st.call_counter += 1
    cc = st.call_counter
    sum_diff = f32(0)
    meas_sum = f32(0)
    sim_in = f32(0)
    sim_wu = f32(0)
    sim_sum = f32(0)
    ny = 0
    for i in range(n):                                    # calibration.cpp:300-313
        if meas[i] > -1:
            ny += 1
            sum_diff = f32(sum_diff + f32(sim[i] - meas[i]))
            meas_sum = f32(meas_sum + meas[i])
            sim_sum = f32(sim_sum + sim[i])
            sim_in = f32(sim_in + inflow[i])
            sim_wu = f32(sim_wu + water_use[i])
    meas_avg = f32(meas_sum / f32(ny))                    # calibration.cpp:315
    crit = abs(float(f32(sum_diff / f32(f32(ny) * meas_avg)))) < 0.01
    up, lo = st.upper, st.lower
    gamma = f32(-98)  # an unset local; see calibration.cpp:409, 437
    if crit:
        gamma = f32(-99)
        st.gamma_cond = 1
        st.calib_status = 1
    elif cc == 1:                                          # calibration.cpp:318-366
        if sum_diff > 0:
            if gamma_old >= up:
                gamma = f32(-99)
            else:
                st.gamma_low = gamma_old
                gamma = up
        else:
            if gamma_old <= lo:
                gamma = f32(-99)
            else:
                st.gamma_high = gamma_old
                gamma = lo
    elif cc == 2:                                          # calibration.cpp:367-448
        if sum_diff > 0:
            if gamma_old == up:
                gamma = gamma_old
                st.gamma_high = up
            elif gamma_old < up:
                st.gamma_low = gamma_old
                if st.gamma_high < 0:
                    gamma = f32(float(gamma_old) * 2.0)
                else:
                    gamma = f32(float(f32(st.gamma_low + st.gamma_high)) / 2.0)
        else:
            if gamma_old == lo:
                gamma = gamma_old
                st.gamma_low = lo
            elif gamma_old > lo:
                if st.gamma_low < 0:
                    gamma = f32(float(gamma_old) / 2.0)
                    st.gamma_high = gamma_old
                else:
                    st.gamma_high = gamma_old
                    gamma = f32(float(f32(st.gamma_low + st.gamma_high)) / 2.0)
    else:                                                  # calibration.cpp:449-523
        if sum_diff > 0:
            if gamma_old == up:
                gamma = f32(-99)
                st.gamma_cond = 2
            elif gamma_old < up:
                st.gamma_low = gamma_old
                if st.gamma_high < 0:
                    gamma = f32(float(gamma_old) * 2.0)
                else:
                    gamma = f32(float(f32(st.gamma_low + st.gamma_high)) / 2.0)
        else:
            if gamma_old == lo:
                gamma = f32(-99)
                st.gamma_cond = 2
            elif gamma_old > lo:
                if st.gamma_low < 0:
                    gamma = f32(float(gamma_old) / 2.0)
                    st.gamma_high = gamma_old
                else:
                    st.gamma_high = gamma_old
                    gamma = f32(float(f32(st.gamma_low + st.gamma_high)) / 2.0)
    adapt = meas_avg                                        # calibration.cpp:549-597
    mean_sim = f32(sim_sum / f32(ny))
    if gamma_old == up and st.gamma_cond == 2:
        adapt = f32(float(meas_avg) * 1.1)
        if mean_sim < adapt:
            st.calib_status = 2
        else:
            st.cell_corr_factor_ind = 99
    if gamma_old == lo and st.gamma_cond == 2:
        adapt = f32(float(meas_avg) * 0.9)
        if mean_sim > adapt:
            st.calib_status = 2
        else:
            st.cell_corr_factor_ind = 99
    if st.cell_corr_factor_ind == 99:
        num = f32(adapt + f32(f32(sim_wu - sim_in) / f32(ny)))
        den = f32(f32(f32(sim_sum + sim_wu) - sim_in) / f32(ny))
        st.cell_corr_factor = f32(num / den)
    if gamma > 0 and float(gamma) < 0.99 * float(lo) and cc > 2:      # calibration.cpp:613-617
        gamma = f32(-99)
    if gamma > 0 and abs(float(f32(f32(st.sum_diff_old / sum_diff) - f32(1)))) < 0.0001:
        gamma = f32(-99)                                               # calibration.cpp:618-623
    st.sum_diff_old = sum_diff
    if gamma < 0 and abs(float(f32(st.cell_corr_factor - f32(1)))) > 0.01:
        st.calib_status = 3
    st.measured_avg_adapt = adapt
    st.sim_runoff_avg = mean_sim
    return gamma


def correction_grid(calib_basin, station_number, annual_pot, simulated, measured, corr):
    """calibration.cpp:730-801 (createCorrectionGrid) in float: the mean of the
    annual grids, the basin's sum of absolute means (in cell order), and for each
    basin cell 1 - sign(mean) * (simulated - measured) / sum, limited to 0.5 .. 1.5.
    Returns (corr updated, mean)."""
    annual = np.asarray(annual_pot, dtype=np.float32)
    ny = annual.shape[0]
    mean = np.zeros(annual.shape[1], dtype=np.float32)
    for y in range(ny):
        mean = (mean + annual[y]).astype(np.float32)
    mean = (mean / np.float32(ny)).astype(np.float32)
    cells = np.flatnonzero(np.asarray(calib_basin) == station_number)
    total = f32(0)
    for i in cells:
        total = f32(total + abs(mean[i]))
    corr = np.array(corr, dtype=np.float64)
    diff = f32(f32(simulated) - f32(measured))
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in cells:
            sg = 1 if mean[i] > 0 else (0 if mean[i] == 0 else -1)   # sign(), calibration.cpp:32-43
            c = float(f32(f32(1) - f32(f32(f32(sg) * diff) / total)))
            if c > 1.5:
                c = 1.5
            if c < 0.5:
                c = 0.5
            corr[i] = c
    return corr, mean


def station_corr_factor(st, gamma, measured, simulated):
    """calibration.cpp:656-728 (writeCorrFactors): the station correction factor
    CFS in float; sets calib_status 4 when it differs from 1."""
    meas_sum = f32(0)
    sim_sum = f32(0)
    for m, s in zip(np.asarray(measured, np.float32), np.asarray(simulated, np.float32)):
        if m > -1:
            meas_sum = f32(meas_sum + m)
            sim_sum = f32(sim_sum + s)
    gamma = f32(gamma)
    if gamma == st.upper and st.cell_corr_factor_ind == 99:   # calibration.cpp:684-687
        meas_sum = f32(float(meas_sum) * 1.1)
    if gamma == st.lower and st.cell_corr_factor_ind == 99:   # calibration.cpp:688-691
        meas_sum = f32(float(meas_sum) * 0.9)
    cfs = f32(meas_sum / sim_sum) if st.cell_corr_factor_ind == 99 else f32(1.0)   # calibration.cpp:694-700
    if (1.0 < cfs < 1.01) or (0.99 < cfs < 1.0):              # calibration.cpp:704-709
        cfs = f32(1.0)
    if cfs > 1.0 or cfs < 1.0:                                 # calibration.cpp:710-711
        st.calib_status = 4
    return cfs


def _upstream_basin(cell, basin_no, inflow9, grid):
    """calib_basins.cpp:302-378 (upstreamBasin), sweep by sweep over all cells."""
    ng = len(grid)
    upstream = downstream = False
    base = 0
    if grid[cell - 1] != 0:
        upstream = True
        base = grid[cell - 1]
    grid[cell - 1] = -1
    while True:
        not_ready = False
        for n in range(ng):
            if grid[n] == -1:
                for j in range(9):
                    grid[n] = -2
                    u = inflow9[n][j]
                    if u != 0:
                        if grid[u - 1] == base:
                            not_ready = True
                            grid[u - 1] = -1
                        else:
                            downstream = True
        if not not_ready:
            break
    for n in range(ng):
        if grid[n] == -2:
            grid[n] = basin_no
    return (1 if upstream else 0) + (2 if downstream else 0)


def calib_basins(inflow9, outflow, station_cells):
    """calib_basins.cpp:25-300 (cbasinClass::prepare) on the station cells in file
    order; returns (grid, used, overlap)."""
    ng = len(outflow)
    grid = [0] * ng
    used, overlap, accepted = [], [], []
    for cn in station_cells:
        if cn <= 1 or cn in accepted:          # calib_basins.cpp:117-131
            used.append(0)
            overlap.append(-1)
            continue
        accepted.append(cn)
        overlap.append(_upstream_basin(cn, len(accepted), inflow9, grid))
        used.append(1)
    for cn in accepted:                        # calib_basins.cpp:270-285
        o = outflow[cn - 1]
        if o != 0 and grid[o - 1] == 0:
            grid[o - 1] = -1
    return np.array(grid, np.int16), used, overlap


def upstream_stations(inflow9, basin_cells, all_stations):
    """upstream_stations.cpp:66-140 (findStations), sweep by sweep; returns one
    list per basin in the reference's order (new entries at the front)."""
    ng = len(inflow9)
    done = [0] * ng
    lists = []
    for cb, cell in enumerate(basin_cells):
        lst = []
        done[cell - 1] = 1
        while True:
            for n in range(ng):
                if done[n] == 1:
                    done[n] = 2
                    for i in range(9):
                        u = inflow9[n][i]
                        if u != 0:
                            st = 0
                            for cb2, c2 in enumerate(basin_cells):
                                if u == c2:
                                    st = cb2 + 1
                                    break
                            if st:
                                done[u - 1] = all_stations
                                lst.insert(0, st)
                            else:
                                done[u - 1] = 1
            if not any(d == 1 for d in done):
                break
        lists.append(lst)
    return lists
