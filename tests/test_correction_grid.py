"""Cell correction factors of a calibration basin (calibGammaClass::
createCorrectionGrid, calibration.cpp:730-801): known answers traced from the
reference's arithmetic, the library against the float restatement on random
grids, and the simulated mean that findNewGamma hands over."""
import ctypes as C

import numpy as np
import pytest

from oracle import calib_oracle

f32 = np.float32


def lib_grid(wg, basin, station, annual, sim, meas, corr):
    lib = wg._abi.load_library()
    basin = np.ascontiguousarray(basin, dtype=np.int16)
    annual = np.ascontiguousarray(annual, dtype=np.float32)
    corr = np.array(corr, dtype=np.float64)
    mean = np.zeros(annual.shape[1], np.float32)
    rc = lib.wg_cell_corr_factor_grid(annual.shape[1], wg._abi.ptr(basin, wg._abi.P_i16), int(station),
                                      annual.shape[0], wg._abi.ptr(annual, wg._abi.P_f), float(sim), float(meas),
                                      wg._abi.ptr(mean, wg._abi.P_f), wg._abi.ptr(corr, wg._abi.P_d))
    assert rc == 0
    return corr, mean


def test_known_answers(wg):
    # basin 1: cells 0-2; cell 3 belongs to basin 2 and keeps its factor
    basin = [1, 1, 1, 2]
    annual = [[2.0, -1.0, 0.0, 5.0], [4.0, -1.0, 0.0, 5.0]]   # means 3, -1, 0 (sum of |mean| 4)
    corr0 = [1.0, 1.0, 1.0, 0.77]
    # simulated exceeds measured by 2: positive cells lose 2/4, negative ones gain
    for impl in ("lib", "oracle"):
        if impl == "lib":
            corr, mean = lib_grid(wg, basin, 1, annual, 10.0, 8.0, corr0)
        else:
            corr, mean = calib_oracle.correction_grid(basin, 1, annual, 10.0, 8.0, corr0)
        assert list(mean) == [3.0, -1.0, 0.0, 5.0]
        assert list(corr) == [0.5, 1.5, 1.0, 0.77], impl
    # a difference of 6 leaves the limits: 1 - 6/4 and 1 + 6/4
    for impl in ("lib", "oracle"):
        corr = (lib_grid(wg, basin, 1, annual, 14.0, 8.0, corr0)[0] if impl == "lib"
                else calib_oracle.correction_grid(basin, 1, annual, 14.0, 8.0, corr0)[0])
        assert list(corr) == [0.5, 1.5, 1.0, 0.77]
    # measured above simulated by 1: 1 + 1/4 and 1 - 1/4
    corr, _ = lib_grid(wg, basin, 1, annual, 7.0, 8.0, corr0)
    assert list(corr) == [1.25, 0.75, 1.0, 0.77]
    # the other basin: one cell, |mean| 5, difference 1 -> 1 - 1/5 in float
    corr, _ = lib_grid(wg, basin, 2, annual, 9.0, 8.0, corr0)
    assert corr[3] == float(f32(1) - f32(1) / f32(5)) and list(corr[:3]) == [1.0, 1.0, 1.0]


def test_float_arithmetic_matches_restatement(wg):
    """The mean and the sum are float, as the reference's Grid<float> (the annual
    files hold floats); the library gives the restatement's bits on grids where
    float and double results differ."""
    rng = np.random.default_rng(3)
    n, ny = 500, 7
    basin = rng.integers(0, 4, n).astype(np.int16)
    annual = (rng.lognormal(0.0, 2.0, (ny, n)) * rng.choice([-1.0, 1.0], (ny, n))).astype(np.float32)
    annual[:, ::17] = 0.0
    corr0 = rng.uniform(0.5, 1.5, n)
    for station in (1, 2, 3):
        for sim, meas in ((120.5, 100.25), (3.0, 9.0), (1e3, 1e-3)):
            a, ma = lib_grid(wg, basin, station, annual, sim, meas, corr0)
            b, mb = calib_oracle.correction_grid(basin, station, annual, sim, meas, corr0)
            assert np.array_equal(ma, mb)
            assert np.array_equal(a, b)
            assert np.all((a[basin == station] >= 0.5) & (a[basin == station] <= 1.5))
            assert np.array_equal(a[basin != station], corr0[basin != station])
    # the float mean differs from the double mean somewhere (the test exercises float)
    dmean = annual.astype(np.float64).mean(axis=0)
    assert np.any(ma.astype(np.float64) != dmean)


def test_search_hands_over_the_simulated_mean(wg):
    """findNewGamma passes simRunoffSum / numberOfYears and the adapted measured
    mean to createCorrectionGrid (calibration.cpp:627)."""
    lib = wg._abi.load_library()
    rng = np.random.default_rng(9)
    for _ in range(50):
        ny = int(rng.integers(3, 12))
        meas = rng.uniform(1.0, 50.0, ny).astype(np.float32)
        meas[rng.random(ny) < 0.2] = -99.0
        if not (meas > -1).any():
            meas[0] = 10.0
        sim = (meas * rng.uniform(0.5, 1.5)).astype(np.float32)
        s = wg._abi.GammaSearch()
        lib.wg_gamma_search_init(C.byref(s))
        st = calib_oracle.GammaSearchState()
        g = lib.wg_find_new_gamma(C.byref(s), C.c_float(1.0), ny, wg._abi.ptr(meas, wg._abi.P_f),
                                  wg._abi.ptr(sim, wg._abi.P_f), None, None)
        go = calib_oracle.find_new_gamma(st, 1.0, meas, sim)
        assert f32(g) == go
        assert f32(s.sim_runoff_avg) == st.sim_runoff_avg
        assert f32(s.measured_avg_adapt) == st.measured_avg_adapt
        ok = meas > -1
        assert abs(float(s.sim_runoff_avg) - float(sim[ok].mean())) <= 1e-5 * float(sim[ok].mean())


def test_bad_arguments_are_refused(wg):
    lib = wg._abi.load_library()
    P_f, P_d, P_i16 = wg._abi.P_f, wg._abi.P_d, wg._abi.P_i16
    basin = np.ones(4, np.int16)
    annual = np.ones((2, 4), np.float32)
    corr = np.ones(4)
    ptr = wg._abi.ptr
    args = [4, ptr(basin, P_i16), 1, 2, ptr(annual, P_f), 1.0, 1.0, None, ptr(corr, P_d)]
    assert lib.wg_cell_corr_factor_grid(*args) == 0
    for k, bad in ((0, 0), (3, 0), (1, None), (4, None), (8, None)):
        a = list(args)
        a[k] = bad
        assert lib.wg_cell_corr_factor_grid(*a) == -1, k  # WG_E_ARG
