"""Checks that pin the CPU oracle itself, independently of the GPU path.

The reference ships no golden vectors for this path, so each check here rests on
something the oracle's own formulas do not restate: a published known-answer
vector, a conservation law, a closed-form special case, a numerical ODE
integration or textbook table values.
"""
import numpy as np
import pytest


def _fresh_regional(wg):
    """The 2,000-cell regional grid without built-up area (the reference drops the
    immediate runoff of built-up area in the snow branch, daily.cpp:1066-1071 and
    1216-1222, which is a real loss the closure would otherwise have to model)."""
    w = wg.World(wg.regional_spec(seed=0))
    w.built_up[:] = 0.0
    return w


def test_philox_known_answers(oracle):
    """Philox4x32-10 against the known-answer vectors published with Random123
    (kat_vectors: zero, all-ones and pi-digit counter/key)."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, out in kat:
        assert oracle.philox(ctr, key) == out


def test_zero_sigma_perturbation_is_identity(oracle, regional_world):
    f = regional_world.forcing(seed=0, day0=0, n_days=1)
    t, p = oracle.perturb_forcing(7, 0, f[0, 0], f[0, 1], 0.0, 0.0)
    assert np.array_equal(t, f[0, 0]) and np.array_equal(p, f[0, 1])


def _vertical_storage(st):
    """Canopy + mean sub-cell snow + soil (mm) per cell."""
    return st["canopy"] + st["snow_elev"][:, 1:].sum(axis=1) / 100.0 + st["soil"]


def test_vertical_closure_one_year(wg, oracle):
    """Per-cell water balance of calcNewDay over 365 days: with CFA = 1 the change of
    canopy + snow + soil storage, after the reference's land-area rescaling
    S * LAF_prev / LAF (daily.cpp:836, 945, 1089), equals precipitation minus
    evapotranspiration, surface runoff and groundwater recharge to < 1e-9 mm/day."""
    w = _fresh_regional(wg)
    n = w.n
    o = oracle.Oracle(w)
    o.set_params(cfa=np.ones(n))
    f = w.forcing(seed=0, day0=0, n_days=365)
    worst = 0.0
    checked = 0
    smax = w.smax.astype(np.float64)
    for d in range(365):
        before = o.get_state()
        o.step_day(f[d, 0], f[d, 1], f[d, 2], f[d, 3], d + 1)
        after = o.get_state()
        out = o.day_outputs()
        laf = before["laf"]
        laf_prev = before["laf_prev"] if before["started"] else laf
        ok = (laf > 0) & (smax > 0)
        s0 = _vertical_storage(before)
        s1 = _vertical_storage(after)
        flux = out["prec"] - out["cell_aet_uncorr"] - out["surface_runoff"] - out["gw_recharge"]
        res = s1[ok] - s0[ok] * laf_prev[ok] / laf[ok] - flux[ok]
        worst = max(worst, float(np.max(np.abs(res))))
        checked += int(ok.sum())
    assert checked > 300 * n
    assert worst < 1e-9, worst


def _warm_day_state(o, n, x, smax):
    st = o.get_state()
    st["canopy"][:] = 0.0
    st["snow"][:] = 0.0
    st["snow_elev"][:] = 0.0
    st["soil"][:] = x * smax
    st["started"] = 0
    o.set_state(st)


def test_hbv_runoff_closed_form(wg, oracle):
    """R = Peff (S/Smax)^gamma (daily.cpp:1122-1127) on a warm wet day:
    S = 0 gives R = 0; S = Smax gives R = Peff, where Peff is read from a run with
    S = 0 (there no runoff and no evapotranspiration leave the soil, so the soil
    gains exactly Peff); S = Smax/2 gives R_a / R_b = 0.5^(gamma_a - gamma_b)."""
    w = _fresh_regional(wg)
    n = w.n
    smax = w.smax.astype(np.float64)
    temp = np.full(n, 25.0, np.float32)
    prec = np.full(n, 20.0, np.float32)  # above P_PCRITGWA, so arid cells recharge too
    sw = np.full(n, 200.0, np.float32)
    lw = np.full(n, 300.0, np.float32)
    gamma_a = w.params[wg.PARAM_INDEX["P_GAMRUN_C"]].copy()
    gamma_b = gamma_a * 0.5 + 0.3

    def run(x, gamma):
        o = oracle.Oracle(w)
        o.set_params(gamma=gamma, cfa=np.ones(n))
        _warm_day_state(o, n, x, smax)
        o.step_day(temp, prec, sw, lw, 180)
        out = o.day_outputs()
        return out["surface_runoff"] + out["gw_recharge"], o.get_state()

    land = (w.contfreq > 0) & (smax > 0)
    r0, s0 = run(0.0, gamma_a)
    assert np.all(r0[land] == 0.0)
    peff = s0["soil"]
    assert np.all(peff[land] > 0) and np.all(peff[land] <= 20.0)
    r1, _ = run(1.0, gamma_a)
    np.testing.assert_allclose(r1[land], peff[land], rtol=1e-13, atol=0)
    rh_a, _ = run(0.5, gamma_a)
    rh_b, _ = run(0.5, gamma_b)
    np.testing.assert_allclose(rh_a[land] / rh_b[land], 0.5 ** (gamma_a[land] - gamma_b[land]),
                               rtol=1e-12)
    np.testing.assert_allclose(rh_a[land], peff[land] * 0.5 ** gamma_a[land], rtol=1e-12)


def test_groundwater_linear_reservoir_vs_rk4(wg, oracle):
    """Humid-cell groundwater (routing.cpp:1968-2039) solves dS/dt = I - k S over one
    day; compare with a fine RK4 integration from the same start, I taken from the
    day's recharge (mm over the land area, routing.cpp:1975)."""
    w = _fresh_regional(wg)
    n = w.n
    o = oracle.Oracle(w)
    f = w.forcing(seed=2, day0=0, n_days=40)
    k = w.params[wg.PARAM_INDEX["P_GWOUTF_C"]]
    humid = (w.aindex != 1) & (w.ldd >= 0) & (w.contfreq > 0)
    worst, n_big = 0.0, 0
    for d in range(40):
        before = o.get_state()
        o.step_day(f[d, 0], f[d, 1], f[d, 2], f[d, 3], d + 1)
        after = o.get_state()
        out = o.day_outputs()
        inflow = out["gw_recharge"] * w.cell_area * (before["laf"] / 100.0) / 1e6
        s = before["groundwater"].copy()
        steps = 400
        h = 1.0 / steps
        for _ in range(steps):
            k1 = inflow - k * s
            k2 = inflow - k * (s + 0.5 * h * k1)
            k3 = inflow - k * (s + 0.5 * h * k2)
            k4 = inflow - k * (s + h * k3)
            s = s + h / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        # the reference zeroes storages below minStorVol = 1e-15 km3 (routing.cpp:1977)
        err = np.abs(after["groundwater"] - s)[humid]
        assert np.all(err <= 1e-10 * np.abs(s[humid]) + 1e-15)
        big = np.abs(s[humid]) > 1e-9
        if big.any():
            worst = max(worst, float(np.max(err[big] / np.abs(s[humid][big]))))
            n_big += int(big.sum())
    assert n_big > 1000
    assert worst < 1e-10, worst


def test_priestley_taylor_against_fao56_tables(wg, oracle):
    """Open-water Priestley-Taylor PET (daily.cpp:476-522) at 20 degC with the slope of
    the saturation vapour pressure curve and the psychrometric constant taken from
    FAO-56 tables 2.4 (0.145 kPa/degC at 20 degC) and 2.2 (0.067 kPa/degC at sea
    level): the oracle's own Delta and gamma must agree to table precision."""
    w = _fresh_regional(wg)
    n = w.n
    o = oracle.Oracle(w)
    t = 20.0
    sw, lw = 200.0, 330.0
    o.step_day(np.full(n, t, np.float32), np.zeros(n, np.float32), np.full(n, sw, np.float32),
               np.full(n, lw, np.float32), 180)
    pet = o.day_outputs()["open_water_pet"]
    lam = 2.501 - 0.002361 * t  # MJ/kg
    eps = np.array([w.lct.emissivity[c - 1] for c in w.land_cover])
    rn = 0.0864 / lam * (sw * (1 - 0.08) + lw) - eps * 4.903e-9 * (t + 273.2) ** 4 / lam
    alpha = np.where(w.aindex == 1, w.params[wg.PARAM_INDEX["P_PTC_ARI"]],
                     w.params[wg.PARAM_INDEX["P_PTC_HUM"]])
    delta, gamma = 0.145, 0.067
    expect = alpha * delta / (delta + gamma) * rn
    land = w.contfreq > 0
    np.testing.assert_allclose(pet[land], expect[land], rtol=5e-3)


def test_perturbation_deviates_are_standard_normal(oracle):
    """The ensemble's Box-Muller deviates: 2 x 100,000 draws (cells x days of one
    member) pass a Kolmogorov-Smirnov test against N(0, 1) at the 1 % level, with
    the pair's two deviates uncorrelated and the 3-sigma tail at its normal share."""
    from scipy import stats
    cells = np.arange(10000, dtype=np.uint32)
    z = np.concatenate([oracle.normal_pairs(12345, cells, day) for day in range(10)])
    for k in (0, 1):
        assert stats.kstest(z[:, k], "norm").pvalue > 0.01
        assert abs(z[:, k].mean()) < 0.01 and abs(z[:, k].std() - 1.0) < 0.01
        assert abs((np.abs(z[:, k]) > 3).mean() - 0.0026998) < 0.0006
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.01
    # the perturbation applies them as P exp(sigma z0), T + sigma z1
    t, p = oracle.perturb_forcing(12345, 3, np.full(4, 10.0, np.float32), np.full(4, 2.0, np.float32), 0.2, 1.0)
    zz = oracle.normal_pairs(12345, np.arange(4), 3)
    np.testing.assert_allclose(p, (2.0 * np.exp(0.2 * zz[:, 0])).astype(np.float32), rtol=1e-6)
    np.testing.assert_allclose(t, (10.0 + zz[:, 1]).astype(np.float32), rtol=1e-6)
