"""Pinning checks for the oracle's storages, outflow laws and state machines.

The reference ships no golden vectors for this path, so every function of the
oracle is held to something its own code does not restate:

* conservation of water in every routing store, per cell and per basin, in every
  clamp branch (spills, dry stores, local lakes drawn below zero, recharge below
  surface water bodies in arid cells);
* the published model equations evaluated independently in numpy on cases chosen
  so that each term is visible: the lake and wetland outflow laws
  Q = k S (S/Smax)^1.5 and ^2.5, the evaporation reduction factors, the linear
  reservoir solution against a Runge-Kutta integration, Manning's velocity on the
  2:1 trapezoidal channel, the reservoir release rules of Hanasaki et al. (2006)
  as WaterGAP modifies them (Doell et al. 2009);
* hand-derived known answers for the LAI growing-season machine (lai.cpp:179-293)
  and for the threshold-elevation rule of the snow sub-cells (daily.cpp:958-976);
* the land-area bookkeeping identity: land + surface water bodies + river area of
  a cell change together (routing.cpp:5034-5188, 5343-5352).

The last test builds mutants of the oracle (one planted mistake each: an exponent
swapped, a factor dropped, a term removed) and shows that the check written for
that arithmetic fails on it.
"""
import math
import os

import numpy as np

import test_water_use_oracle as wu
import pytest

from oracle import wg_oracle

KSWBGW = 10.0        # mm/d of focused recharge below surface water in arid cells (routing.cpp:1662)
EVAPO_RED = 3.32193  # routing.cpp:124-129
EVAPO_RED_RES = 2.81383


def PI(wg, name):
    return wg.PARAM_INDEX[name]


# ---------------------------------------------------------------------------
# worlds and states
# ---------------------------------------------------------------------------

def chain_world(wg, seed=3):
    """A 1 x 8 strip: cells 0 and 7 are headwaters, cell 2 the outlet
    (0 -> 1 -> 2 <- 3 <- 4 <- 5 <- 6 <- 7).  No surface water body, no built-up
    area, humid, CFA = 1; the checks switch on what they need."""
    w = wg.World(wg.regional_spec(seed=seed, nrows=1, ncols=8))
    for f in ("glo_lake", "loc_lake", "glo_wetland", "loc_wetland", "lake_area", "reservoir_area", "glo_res",
              "stor_cap", "mean_outflow"):
        getattr(w, f)[:] = 0.0
    w.built_up[:] = 0.0
    w.contfreq[:] = 100.0
    w.aindex[:] = 0
    w.params[PI(wg, "P_CFA")] = 1.0
    w.params[PI(wg, "M_PREC")] = 1.0
    assert list(w.downstream) == [2, 3, 0, 3, 4, 5, 6, 7]
    return w


def dry_state(o):
    """No water on land (no canopy, snow, soil or groundwater), so that no runoff
    reaches the routing stores."""
    st = o.get_state()
    for k in ("canopy", "soil", "snow", "snow_elev", "storage_transfer", "groundwater"):
        st[k][:] = 0.0
    st["started"] = 0
    return st


def day(n, temp=20.0, prec=0.0, sw=0.0, lw=0.0):
    """Uniform forcing.  With SW = LW = 0 the net radiation is negative, so the
    potential evaporation of land and open water is 0 (daily.cpp:497-520)."""
    return [np.full(n, v, np.float32) for v in (temp, prec, sw, lw)]


def run_one_day(wg, w, st, forcing, doy, variant="wg"):
    o = wg_oracle.Oracle(w, variant)
    o.set_state(st)
    o.step_day(*forcing, doy)
    return o.get_state(), o.day_outputs()


def lake_depth(wg, w, c):
    return w.params[PI(wg, "P_LAK_D"), c] * 0.001


def wetl_depth(wg, w, c):
    return w.params[PI(wg, "P_WET_D"), c] * 0.001


def red_factor(s, smax, denom, expo):
    """Evaporation reduction factor 1 - (|S - Smax| / denom)^expo, clamped to [0, 1]."""
    return min(1.0, max(0.0, 1.0 - math.pow(abs(s - smax) / denom, expo)))


def close(a, b, rtol=1e-12, atol=0.0):
    assert abs(a - b) <= rtol * max(abs(a), abs(b)) + atol, (a, b, (a - b) / max(abs(b), 1e-300))


def rk4_linear(s0, inflow, k, steps=2000):
    """Integrates dS/dt = inflow - k S over one day."""
    h = 1.0 / steps
    s = s0
    for _ in range(steps):
        k1 = inflow - k * s
        k2 = inflow - k * (s + 0.5 * h * k1)
        k3 = inflow - k * (s + 0.5 * h * k2)
        k4 = inflow - k * (s + h * k3)
        s += h / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    return s


# ---------------------------------------------------------------------------
# outflow laws of the local lake and the local wetland
# ---------------------------------------------------------------------------

def check_local_lake(wg, variant="wg", arid=False):
    """Local lake on a headwater cell with no local runoff: the day's evaporation
    PET_ow * rf over the lake area is taken first, then the outflow
    k S(t-1) (S(t-1)/Smax)^1.5 (routing.cpp:2318-2490; Doell et al. 2003 eq. 4),
    and the next day's reduction factor is 1 - (|S - Smax| / (2 Smax))^(3.32193 e).
    In an arid cell the lake also loses 10 mm/d over its reduced area to the
    groundwater."""
    w = chain_world(wg)
    c = 0
    w.loc_lake[c] = 5.0
    if arid:
        w.aindex[c] = 1
    area = w.cell_area[c]
    smax = 5.0 / 100.0 * area * lake_depth(wg, w, c)
    k = w.params[PI(wg, "P_SWOUTF_C"), c]
    e = w.params[PI(wg, "M_EVAREDEX"), c]
    o = wg_oracle.Oracle(w, variant)
    st = dry_state(o)
    s0, rf = smax / 4.0, 0.8
    st["loc_lake"][c] = s0
    st["red_loc_lake"][c] = rf
    after, out = run_one_day(wg, w, st, day(w.n, 20.0, 0.0, 200.0, 330.0), 180, variant)
    pet = out["open_water_pet"][c]
    assert pet > 1.0
    evap = pet * rf * area / 1e6 * 5.0 / 100.0
    gwr = KSWBGW * rf * area / 1e6 * 5.0 / 100.0 if arid else 0.0
    q = k * s0 * (s0 / smax) ** 1.5
    assert s0 - evap - gwr > q
    s1 = s0 - evap - gwr - q
    close(after["loc_lake"][c], s1)
    close(after["red_loc_lake"][c], red_factor(s1, smax, 2.0 * smax, EVAPO_RED * e))
    # a full lake without evaporation loses k Smax (and the recharge of an arid cell)
    st["loc_lake"][c] = smax
    st["red_loc_lake"][c] = 1.0
    after, _ = run_one_day(wg, w, st, day(w.n), 180, variant)
    s1 = smax - gwr / rf - k * smax
    close(after["loc_lake"][c], s1)
    close(after["red_loc_lake"][c], red_factor(s1, smax, 2.0 * smax, EVAPO_RED * e))


def check_local_wetland(wg, variant="wg"):
    """Local wetland: evaporation first, then the outflow k S (S/Smax)^2.5 of the
    storage after evaporation (routing.cpp:2495-2617), and the reduction factor
    1 - (|S - Smax| / Smax)^(3.32193 e)."""
    w = chain_world(wg)
    c = 7
    w.loc_wetland[c] = 4.0
    area = w.cell_area[c]
    smax = 4.0 / 100.0 * area * wetl_depth(wg, w, c)
    k = w.params[PI(wg, "P_SWOUTF_C"), c]
    e = w.params[PI(wg, "M_EVAREDEX"), c]
    o = wg_oracle.Oracle(w, variant)
    st = dry_state(o)
    s0, rf = smax / 4.0, 0.7
    st["loc_wetl"][c] = s0
    st["red_loc_wetl"][c] = rf
    after, out = run_one_day(wg, w, st, day(w.n, 20.0, 0.0, 200.0, 330.0), 180, variant)
    evap = out["open_water_pet"][c] * rf * area / 1e6 * 4.0 / 100.0
    sp = s0 - evap
    s1 = sp - k * sp * (sp / smax) ** 2.5
    close(after["loc_wetl"][c], s1)
    close(after["red_loc_wetl"][c], red_factor(s1, smax, smax, EVAPO_RED * e))


def test_local_lake_outflow_law(wg):
    check_local_lake(wg)


def test_local_lake_arid_recharge(wg):
    check_local_lake(wg, arid=True)


def test_local_wetland_outflow_law(wg):
    check_local_wetland(wg)


# ---------------------------------------------------------------------------
# linear reservoirs: global lake, global wetland, river
# ---------------------------------------------------------------------------

def check_global_lake(wg, variant="wg"):
    """Global lake fed by an upstream river (cell 0 -> lake in cell 1): the storage
    follows dS/dt = I - E - k S over the day (routing.cpp:2630-2804), with I the
    upstream transported volume and E = PET_ow rf A_lake; compared with a Runge-Kutta
    integration.  The reduction factor divides by 2 Smax."""
    w = chain_world(wg)
    c = 1
    w.lake_area[c] = 0.08 * w.cell_area[c]
    w.glo_lake[c] = 8.0
    smax = w.lake_area[c] * lake_depth(wg, w, c)
    k = w.params[PI(wg, "P_SWOUTF_C"), c]
    e = w.params[PI(wg, "M_EVAREDEX"), c]
    o = wg_oracle.Oracle(w, variant)
    st = dry_state(o)
    st["river"][0] *= 40.0  # a full upstream river delivers the lake's inflow
    s0, rf = 0.6 * smax, 0.9
    st["glo_lake"][c] = s0
    st["red_glo_lake"][c] = rf
    after, out = run_one_day(wg, w, st, day(w.n, 20.0, 0.0, 200.0, 330.0), 180, variant)
    inflow = out["discharge"][0]
    evap = out["open_water_pet"][c] * rf * w.lake_area[c] / 1e6
    assert inflow > 10 * evap > 0
    s1 = rk4_linear(s0, inflow - evap, k)
    close(after["glo_lake"][c], s1, rtol=1e-11)
    close(after["red_glo_lake"][c], red_factor(after["glo_lake"][c], smax, 2.0 * smax, EVAPO_RED * e))


def check_global_wetland(wg, variant="wg"):
    """Global wetland on a headwater cell: dS/dt = P - E - k S with P = E = 0 gives
    S0 e^-k, and with evaporation S0 e^-k - (E/k)(1 - e^-k) (routing.cpp:3178-3297);
    the reduction factor divides by Smax."""
    w = chain_world(wg)
    c = 7
    w.glo_wetland[c] = 6.0
    area = w.cell_area[c]
    smax = 6.0 / 100.0 * area * wetl_depth(wg, w, c)
    k = w.params[PI(wg, "P_SWOUTF_C"), c]
    e = w.params[PI(wg, "M_EVAREDEX"), c]
    o = wg_oracle.Oracle(w, variant)
    st = dry_state(o)
    s0, rf = 0.5 * smax, 0.6
    st["glo_wetl"][c] = s0
    st["red_glo_wetl"][c] = rf
    after, _ = run_one_day(wg, w, st, day(w.n), 180, variant)
    close(after["glo_wetl"][c], s0 * math.exp(-k))
    close(after["red_glo_wetl"][c], red_factor(s0 * math.exp(-k), smax, smax, EVAPO_RED * e))
    after, out = run_one_day(wg, w, st, day(w.n, 20.0, 0.0, 200.0, 330.0), 180, variant)
    evap = out["open_water_pet"][c] * rf * area / 1e6 * 6.0 / 100.0
    close(after["glo_wetl"][c], rk4_linear(s0, -evap, k), rtol=1e-11)


def river_geometry(w, c, m_rivrgh):
    """Hydraulic geometry of routing.cpp:476-507 (Allen et al. 1994): bankfull width
    and depth from bankfull flow, bottom width of the 2:1 trapezoid, river length
    scaled by the continental fraction, storage at bankfull."""
    bf = max(float(w.bankfull_flow[c]), 0.05)
    width_bf = 2.71 * bf ** 0.557
    depth_bf = 0.349 * bf ** 0.341
    bottom = width_bf - 4.0 * depth_bf
    length = float(w.river_length[c]) * w.contfreq[c] / 100.0
    smax = length * 0.5 * depth_bf / 1000.0 * (bottom / 1000.0 + width_bf / 1000.0)
    return length, bottom, width_bf, smax


def manning_velocity_km_per_day(stor_km3, length_km, bottom_m, slope, rough, mult):
    """Manning's formula on a trapezoid with side slopes 1:2 (horizontal 2 per
    vertical 1): wet area A = B d + 2 d^2 solved for the depth d, wetted perimeter
    B + 2 d sqrt(5), v = (1/n) R^(2/3) s^(1/2) in m/s, floor 1e-5 km/d."""
    area_m2 = stor_km3 / length_km * 1e6
    d = (-bottom_m + math.sqrt(bottom_m * bottom_m + 8.0 * area_m2)) / 4.0
    r = area_m2 / (bottom_m + 2.0 * d * math.sqrt(5.0))
    v = r ** (2.0 / 3.0) * math.sqrt(slope) / (rough * mult) * 86.4
    return max(v, 1e-5)


def check_river(wg, variant="wg"):
    """River of a headwater cell: a linear reservoir with K = v/L from Manning's
    velocity of the day's starting storage (routing.cpp:3402-3586, 7309-7341).
    Without inflow S1 = S0 e^-K and the transported volume is S0 - S1; with
    evaporation from the river area fraction dS/dt = -E - K S.  The next day's
    reduction factor is 1 - (|S - Smax| / Smax)^(3.32193 e) and the river area
    fraction rf L w / A with the water-surface width B + 4 d."""
    w = chain_world(wg)
    c = 0
    mult = w.params[PI(wg, "M_RIVRGH_C"), c]
    e = w.params[PI(wg, "M_EVAREDEX"), c]
    length, bottom, width_bf, smax = river_geometry(w, c, mult)
    o = wg_oracle.Oracle(w, variant)
    st = dry_state(o)
    s0 = 0.7 * smax
    st["river"][c] = s0
    after, out = run_one_day(wg, w, st, day(w.n), 180, variant)
    v = manning_velocity_km_per_day(s0, length, bottom, float(w.river_slope[c]), float(w.roughness[c]), mult)
    kk = v / length
    s1 = s0 * math.exp(-kk)
    close(after["river"][c], s1)
    close(out["discharge"][c], s0 - s1, rtol=1e-11)
    rf1 = red_factor(s1, smax, smax, EVAPO_RED * e)
    close(after["red_river"][c], rf1)
    depth_km = (-bottom / 1000.0 + math.sqrt((bottom / 1000.0) ** 2 + 8.0 * s1 / length)) / 4.0
    width_km = min(bottom / 1000.0 + 4.0 * depth_km, width_bf / 1000.0)
    close(after["river_frac_next"][c], rf1 * length * width_km * 100.0 / w.cell_area[c], rtol=1e-11)
    # evaporation from the river's area fraction of the cell (riverEvapoOpt = 1)
    raf = st["river_frac_next"][c]
    after, out = run_one_day(wg, w, st, day(w.n, 20.0, 0.0, 200.0, 330.0), 180, variant)
    evap = out["open_water_pet"][c] * raf / 100.0 * w.cell_area[c] / 1e6
    assert evap > 0
    close(after["river"][c], rk4_linear(s0, -evap, kk), rtol=1e-11)


def test_global_lake_linear_reservoir(wg):
    check_global_lake(wg)


def test_global_wetland_linear_reservoir(wg):
    check_global_wetland(wg)


def test_river_manning_linear_reservoir(wg):
    check_river(wg)


# ---------------------------------------------------------------------------
# reservoirs (routing.cpp:2807-3082)
# ---------------------------------------------------------------------------

def reservoir_world(wg, c_ratio, res_type=2):
    w = chain_world(wg)
    c = 0
    area = w.cell_area[c]
    w.reservoir_area[c] = 0.05 * area
    w.glo_res[c] = 5.0
    w.stor_cap[c] = 0.5
    # c_ratio = capacity / mean annual inflow volume
    w.mean_outflow[c] = w.stor_cap[c] / c_ratio / 31536000.0 * 1e9
    w.res_type[c] = res_type
    w.start_month[c] = 7  # operational year starts on 1 July = julian day 182
    return w, c


def check_reservoir(wg, variant="wg"):
    """Reservoir release: on the first day of the operational year the release
    coefficient becomes K = S / (0.85 cap), or 0.1 below 10 % of capacity; a
    reservoir with capacity ratio c >= 0.5 releases K Q_mean (m3/s), one with c < 0.5
    blends 4c^2 K Q_mean with (1 - 4c^2) times the day's inflow; below 10 % of
    capacity only a tenth of the release leaves.  The reduction factor has the
    exponent 2.81383 over the capacity."""
    # c >= 0.5, first day of the operational year
    w, c = reservoir_world(wg, c_ratio=0.8)
    cap, q = w.stor_cap[c], w.mean_outflow[c]
    o = wg_oracle.Oracle(w, variant)
    st = dry_state(o)
    s0 = 0.6 * cap
    st["reservoir"][c] = s0
    st["k_release"][c] = 0.3
    after, _ = run_one_day(wg, w, st, day(w.n), 182, variant)
    kr = s0 / (0.85 * cap)
    close(after["k_release"][c], kr)
    s1 = s0 - kr * q * 86400.0 / 1e9
    close(after["reservoir"][c], s1)
    close(after["red_res"][c], red_factor(s1, cap, cap, EVAPO_RED_RES))
    # any other day keeps the coefficient
    after, _ = run_one_day(wg, w, st, day(w.n), 181, variant)
    close(after["k_release"][c], 0.3)
    close(after["reservoir"][c], s0 - 0.3 * q * 86400.0 / 1e9)
    # below 10 % of capacity: K = 0.1 on the first day and a tenth of the release
    st["reservoir"][c] = 0.05 * cap
    after, _ = run_one_day(wg, w, st, day(w.n), 182, variant)
    close(after["k_release"][c], 0.1)
    close(after["reservoir"][c], 0.05 * cap - 0.1 * (0.1 * q) * 86400.0 / 1e9)
    # c < 0.5 without inflow: 4 c^2 K Q_mean
    w, c = reservoir_world(wg, c_ratio=0.2)
    cap, q = w.stor_cap[c], w.mean_outflow[c]
    cr = cap / (q * 31536000.0 / 1e9)
    o = wg_oracle.Oracle(w, variant)
    st = dry_state(o)
    st["reservoir"][c] = 0.6 * cap
    st["k_release"][c] = 0.3
    after, _ = run_one_day(wg, w, st, day(w.n), 181, variant)
    close(after["reservoir"][c], 0.6 * cap - 4.0 * cr * cr * 0.3 * q * 86400.0 / 1e9)


def check_irrigation_reservoir(wg, variant="wg"):
    """Irrigation reservoir (type 1): the mean demand D is the cell's mean net
    surface-water use plus that of up to 5 downstream cells (routing.cpp:360-403,
    m3/yr / 31536000 s); with no daily use the provisional release is Q_mean / 2
    when D >= Q_mean / 2, else Q_mean - D (routing.cpp:2990-2996).  The reference's
    walk tests the reservoir area one cell past the downstream cell
    (G_reservoir_area_full[downstreamCell] with a 1-based number, routing.cpp:372):
    a reservoir two cells down stops the walk after the first."""
    # strip: cell 0 -> cell 1 -> cell 2 (outlet); use on cells 0, 1, 2
    cases = []
    for share, stop_early in ((0.75, False), (0.2, False), (0.75, True)):
        w, c = reservoir_world(wg, c_ratio=0.8, res_type=1)
        q = w.mean_outflow[c]
        nus = np.zeros(w.n)
        nus[[0, 1, 2]] = np.array([0.2, 0.3, 0.5]) * share * q * 31536000.0
        w.mean_nus = nus
        if stop_early:
            w.reservoir_area[2] = 0.01 * w.cell_area[2]
            w.glo_res[2] = 1.0
            w.stor_cap[2] = 0.01
            w.mean_outflow[2] = 1.0
            w.res_type[2] = 2
            w.start_month[2] = 1
        demand = (nus[0] if stop_early else nus[0] + nus[1] + nus[2]) / 31536000.0
        prov = q / 2.0 if demand >= 0.5 * q else q - demand
        o = wg_oracle.Oracle(w, variant)
        st = dry_state(o)
        s0 = 0.6 * w.stor_cap[c]
        st["reservoir"][c] = s0
        st["k_release"][c] = 0.3
        after, _ = run_one_day(wg, w, st, day(w.n), 181, variant)
        close(after["reservoir"][c], s0 - 0.3 * prov * 86400.0 / 1e9)
        cases.append(prov / q)
    # D = 0.75 Q, 0.2 Q, and 0.15 Q when the walk stops after the reservoir cell
    np.testing.assert_allclose(cases, [0.5, 0.8, 0.85], rtol=1e-12)


def test_reservoir_release_rules(wg):
    check_reservoir(wg)


def test_irrigation_reservoir_demand(wg):
    check_irrigation_reservoir(wg)


# ---------------------------------------------------------------------------
# conservation of water in the routing stores
# ---------------------------------------------------------------------------

STORES = ("groundwater", "loc_lake", "loc_wetl", "glo_lake", "glo_wetl", "reservoir", "river")


def conservation_world(wg):
    """The edge-case regional world (lakes every 20 cells, reservoirs every 25,
    land-area fractions near 0) with every surface water body on a good share of
    cells, arid and humid cells and inland sinks."""
    w = wg.World(wg.regional_spec(seed=5, nrows=30, ncols=40, edge_cases=True, glolake_every=20, res_every=25))
    w.built_up[:] = 0.0
    return w


def random_state(o, w, rng):
    """Storages anywhere in their range, beyond it (spills) and below it (a local
    lake can be drawn to -Smax), snow that melts on a warm day and soils above
    capacity, so that water enters the routing stores from land."""
    st = o.get_state()
    n = w.n
    for k in STORES:
        x = st[k].copy()
        st[k] = x * rng.choice([0.0, 0.3, 0.9, 1.0, 1.3], size=n)
    st["loc_lake"] = np.where(rng.random(n) < 0.1, -0.5 * np.abs(st["loc_lake"]), st["loc_lake"])
    st["groundwater"] = rng.uniform(0.0, 0.05, n) * w.cell_area / 1e3
    st["snow_elev"][:, 1:] = np.where(rng.random((n, 1)) < 0.5, rng.uniform(0.0, 30.0, (n, 100)), 0.0)
    st["snow"] = st["snow_elev"][:, 1:].mean(axis=1)
    st["soil"] = w.smax.astype(np.float64) * rng.uniform(0.5, 1.4, n)
    st["started"] = 0
    return st


def volume_in_from_land(before, out, w):
    """Water entering the routing stores from the land of each cell (km3):
    surface runoff and groundwater recharge over the land area
    (routing.cpp:1885-1891, 1975), or the land storages handed over when the land
    fraction is 0 (daily.cpp:825-832, 1080-1090)."""
    laf = before["laf"]
    land = (out["surface_runoff"] + out["gw_recharge"]) * w.cell_area / 1e6 * laf / 100.0
    return np.where(laf > 0, land, 0.0)


def check_conservation(wg, variant="wg", days=6):
    """With no precipitation and no evaporation, per cell and per basin: the change
    of all seven routing storages equals the water from land plus the upstream
    inflow minus the transported volume, in every branch the random states reach."""
    w = conservation_world(wg)
    n = w.n
    o = wg_oracle.Oracle(w, variant)
    rng = np.random.default_rng(11)
    o.set_state(random_state(o, w, rng))
    down = w.downstream.astype(np.int64) - 1
    worst = 0.0
    for d in range(days):
        before = o.get_state()
        o.step_day(*day(n, 15.0 + 3 * d), 150 + d)
        after = o.get_state()
        out = o.day_outputs()
        ds = sum(after[k] - before[k] for k in STORES)
        tv = out["discharge"]
        up = np.zeros(n)
        np.add.at(up, down[down >= 0], tv[down >= 0])
        zero_laf = before["laf"] <= 0
        land = volume_in_from_land(before, out, w)
        # cells without land hand the transfer storage to the routing
        land[zero_laf] = (after["storage_transfer"] * w.cell_area / 1e6 * before["laf_prev"] / 100.0)[zero_laf]
        if d == 0:
            land[zero_laf] = 0.0
        cont = w.contfreq >= 1e-6
        res = (ds - land - up + tv)[cont]
        scale = sum(np.abs(before[k]) + np.abs(after[k]) for k in STORES)[cont] + np.abs(tv[cont]) + np.abs(up[cont])
        assert np.all(np.abs(res) <= 1e-13 * scale + 1e-14), (d, float(np.max(np.abs(res))))
        big = scale > 1e-3
        worst = max(worst, float(np.max(np.abs(res[big]) / scale[big])))
        # basin: storage change plus outlet discharge equals the water from land
        for b, outlet in enumerate(w.basin_outlet):
            cells = (w.basin == b) & cont
            close(float(ds[cells].sum() + tv[outlet]), float(land[cells].sum()), rtol=1e-11, atol=1e-13)
    return worst


def test_routing_conservation_per_cell_and_basin(wg):
    # beyond the absolute floor (storages below 1e-15 km3 are zeroed), the residual
    # is rounding of the largest volumes
    assert check_conservation(wg) < 1e-13


# ---------------------------------------------------------------------------
# land-area bookkeeping
# ---------------------------------------------------------------------------

def check_land_area_identity(wg, variant="wg", days=120):
    """routing.cpp:5173-5178 with 5343-5352: the land-area fraction (%) changes by
    the opposite of the change of the surface water body fraction and of the river
    area fraction, so LAF + 100 fswb + 100 f_river is the same before and after a
    day wherever the land fraction does not hit its floor of 0."""
    w = conservation_world(wg)
    n = w.n
    o = wg_oracle.Oracle(w, variant)
    f = w.forcing(seed=3, day0=0, n_days=days)
    checked = 0
    for d in range(days):
        before = o.get_state()
        o.step_day(f[d, 0], f[d, 1], f[d, 2], f[d, 3], d + 1)
        after = o.get_state()
        tot0 = before["laf"] + 100.0 * before["fswb_frac_next"] + 100.0 * before["river_frac_next"]
        tot1 = after["laf"] + 100.0 * after["fswb_frac_next"] + 100.0 * after["river_frac_next"]
        ok = (after["laf"] > 0) & (w.contfreq >= 1e-6)
        assert np.all(np.abs(tot1 - tot0)[ok] <= 1e-12 * np.maximum(np.abs(tot0[ok]), 1.0)), d
        assert np.array_equal(after["laf_prev"], before["laf"])
        checked += int(ok.sum())
    assert checked > 0.9 * n * days


def test_land_area_identity(wg):
    check_land_area_identity(wg)


# ---------------------------------------------------------------------------
# LAI growing-season machine (lai.cpp:179-293), hand-derived known answers
# ---------------------------------------------------------------------------

def lai_world(wg, lct):
    w = chain_world(wg)
    w.land_cover[:] = lct
    w.params[PI(wg, "M_LAI")] = 1.0
    return w


def lai_limits(w, c):
    """LAImax = M_LAI * LAI(lct) and LAImin = a + b LAImax, in float as lai.cpp:78-89
    and 152-154 compute them."""
    lct = int(w.land_cover[c])
    lmax = np.float32(w.params[14, c] * w.lct.lai[lct - 1])
    decid = np.float32(w.lct.decid_frac[lct - 1])
    a = np.float32(0.1 * float(decid))
    b = np.float32((np.float32(1.0) - decid) * np.float32(w.lct.evergreen_red[lct - 1]))
    lmin = np.float32(a + b * lmax)
    return float(lmin), float(lmax)


def run_lai(wg, w, c, days_spec, variant="wg", arid=False):
    """Runs the listed (temperature, precipitation) days on a land cell without
    potential evaporation and returns, per day, (days_since_start, status,
    precipitation sum, LAI).  The canopy starts empty and the rain fills it to its
    capacity P_MCWH * LAI each day (no evaporation empties it), so the canopy
    storage reads out the day's LAI."""
    if arid:
        w.aindex[c] = 1
    o = wg_oracle.Oracle(w, variant)
    st = dry_state(o)
    o.set_state(st)
    mcwh = w.params[PI(wg, "P_MCWH"), c]
    rows = []
    for k, (t, p) in enumerate(days_spec):
        o.step_day(*day(w.n, t, p), 100 + k)
        s = o.get_state()
        rows.append((int(s["lai_days"][c]), int(s["lai_status"][c]), float(s["lai_precsum"][c]),
                     float(s["canopy"][c]) / mcwh))
    return rows


def check_lai(wg, variant="wg"):
    """A deciduous class (lct 4 in the table) through sprouting, full cover and
    senescence; the evergreen rule of classes 1-2; the arid dry-day rule; the reset
    when the precipitation threshold is not met."""
    lct = 4
    w = lai_world(wg, lct)
    c = 0
    d0 = int(w.lct.initial_days[lct - 1])
    assert d0 >= 4
    lmin, lmax = lai_limits(w, c)
    frac = lambda f: lmin + (lmax - lmin) * f  # noqa: E731
    # growing conditions (T > 8) with 10 mm/d: d0 days of waiting, then the ramp of
    # 30 days, then full cover held at days = d0 + 30
    spec = [(20.0, 10.0)] * (d0 + 33)
    rows = run_lai(wg, w, c, spec, variant)
    for k in range(d0):
        assert rows[k][:3] == (k + 1, 0, 10.0 * (k + 1)), (k, rows[k])
        close(rows[k][3], lmin, rtol=1e-14)
    for j in range(1, 31):
        days, status, psum, lai = rows[d0 - 1 + j]
        assert days == d0 + j and psum == 10.0 * (d0 + j)
        assert status == (1 if j == 30 else 0)
        close(lai, frac(j / 30.0), rtol=1e-14)
    for k in range(d0 + 30, d0 + 33):
        assert rows[k][:2] == (d0 + 30, 1)
        close(rows[k][3], lmax, rtol=1e-14)
    # then cold days (T <= 8): days count down from d0 + 30 at full cover; from 30
    # down the LAI falls linearly to LAImin at 0, where status and the sum reset
    spec2 = spec + [(5.0, 10.0)] * (d0 + 31)
    rows = run_lai(wg, w, c, spec2, variant)
    base = d0 + 33
    for j in range(d0 + 31):
        days, status, psum, lai = rows[base + j]
        expect_days = d0 + 30 - (j + 1)
        assert days == max(expect_days, 0), (j, days)
        if expect_days >= 30:
            close(lai, lmax, rtol=1e-14)
        else:
            close(lai, lmax - (lmax - lmin) * (30 - max(expect_days, 0)) / 30.0, rtol=1e-14)
        assert status == (0 if expect_days <= 0 else 1)
        if expect_days == 0:
            assert psum == 0.0
    # not enough rain: after the waiting days the count stays at d0 and the LAI at
    # LAImin while the sum stays at or below 40 mm
    rows = run_lai(wg, w, c, [(20.0, 1.0)] * (d0 + 10), variant)
    for k in range(d0, d0 + 10):
        assert rows[k][0] == d0 and rows[k][1] == 0
        close(rows[k][3], lmin, rtol=1e-14)
    # evergreen classes (1, 2): in senescence the status drops to 0 at once
    w2 = lai_world(wg, 1)
    d1 = int(w2.lct.initial_days[0])
    st2 = run_lai(wg, w2, c, [(20.0, 10.0)] * (d1 + 31) + [(5.0, 0.0)] * (d1 + 1) + [(20.0, 0.0)], variant)
    sen = st2[d1 + 31 + d1]  # the first cold day with days <= 30
    assert sen[0] == 29 and sen[1] == 1  # the cold branch keeps the status
    assert st2[-1][0] == 28 and st2[-1][1] == 0  # the growing branch drops it for classes 1-2
    # arid cell at full cover: a dry day (P < 0.5 mm) counts down, a wet one resets
    w3 = lai_world(wg, lct)
    rows = run_lai(wg, w3, c, [(20.0, 10.0)] * (d0 + 31) + [(20.0, 0.0)] * 3 + [(20.0, 1.0)], variant, arid=True)
    assert [r[0] for r in rows[d0 + 31:]] == [d0 + 29, d0 + 28, d0 + 27, d0 + 30]


def test_lai_growing_season_known_answers(wg):
    check_lai(wg)


# ---------------------------------------------------------------------------
# threshold elevation of the snow sub-cells (daily.cpp:958-976)
# ---------------------------------------------------------------------------

def check_threshold_elevation(wg, variant="wg"):
    """Sub-cells with more than 1000 mm of snow above the first such sub-cell take
    that sub-cell's temperature instead of their own, colder one; the first takes its
    own.  Warm day, no precipitation, no evaporation: each sub-cell melts
    DDF (T_eff - T_melt), capped at its storage."""
    w = chain_world(wg)
    c = 0
    z0 = 500
    elev = np.zeros(101, np.int16)
    elev[0] = z0
    elev[1:] = z0 + 20 * np.arange(1, 101) - 1000  # 1-sub-cell steps of 20 m, sorted ascending
    elev[1:] = np.maximum(elev[1:], 1)
    w.elevation[c] = elev
    w.params[PI(wg, "P_T_GRADNT"), c] = 0.006
    o = wg_oracle.Oracle(w, variant)
    st = dry_state(o)
    rng = np.random.default_rng(4)
    snow = rng.uniform(0.0, 900.0, 100)
    big = np.array([10, 11, 30, 55, 56, 57, 90])  # 0-based sub-cells above 1000 mm
    snow[big] = rng.uniform(1001.0, 1500.0, len(big))
    st["snow_elev"][c, 1:] = snow
    st["snow"][c] = snow.mean()
    temp = 12.0
    after, _ = run_one_day(wg, w, st, day(w.n, temp), 180, variant)
    lct = int(w.land_cover[c])
    ddf = w.params[PI(wg, "M_DEGDAY_F"), c] * w.lct.ddf[lct - 1]
    t_melt = w.params[PI(wg, "P_T_SNOWMT"), c]
    grad = w.params[PI(wg, "P_T_GRADNT"), c]
    z_thresh = None
    expect = np.empty(100)
    for j in range(100):
        z = int(elev[j + 1])
        if snow[j] > 1000.0:
            if z_thresh is None:
                z_thresh = z
            else:
                z = z_thresh
        t = temp - (z - z0) * grad
        melt = ddf * (t - t_melt) if t > t_melt else 0.0
        expect[j] = max(snow[j] - melt, 0.0)
    np.testing.assert_allclose(after["snow_elev"][c, 1:], expect, rtol=1e-13, atol=1e-12)
    # the rule matters: the upper big sub-cells melt more than their own temperature allows
    own = np.array([max(snow[j] - ddf * max(temp - (int(elev[j + 1]) - z0) * grad - t_melt, 0.0), 0.0)
                    for j in range(100)])
    assert np.all(expect[big[2:]] < own[big[2:]])


def test_threshold_elevation_rule(wg):
    check_threshold_elevation(wg)


# ---------------------------------------------------------------------------
# potential cell runoff (routing.cpp:1962-1970, 2045-2053, 2180-2187, 3372-3379)
# ---------------------------------------------------------------------------

def check_pot_cell_runoff(wg, variant="wg"):
    """The cell's potential runoff is the volume it generates: its surface runoff
    on the land fraction, plus the groundwater it discharges (recharge on the land
    fraction minus the storage change, whichever groundwater equation applied), plus
    the open-water balance (P - PET) * CFA over its lake, reservoir and wetland area.
    Summed over 40 days this holds in every humid cell, every inland sink and every
    (semi-)arid cell without surface water bodies (their groundwater recharge from
    surface water does not show in the day's outputs), with CFA drawn per cell."""
    w = wg.World(wg.global_spec(seed=2, n_land=3000, n_outlets=30, n_sinks=15))
    arid = np.flatnonzero((w.aindex == 1) & (w.ldd >= 0))
    dry = arid[::2]
    for name in ("loc_lake", "loc_wetland", "glo_wetland", "lake_area", "reservoir_area", "glo_lake",
                 "glo_res", "stor_cap"):
        getattr(w, name)[dry] = 0
    cfa = w.params[PI(wg, "P_CFA")]
    cfa[:] = np.random.default_rng(0).uniform(0.7, 1.3, w.n)
    o = wg_oracle.Oracle(w, variant)
    n_days = 40
    f = w.forcing(seed=1, day0=0, n_days=n_days)
    gw0 = o.get_state()["groundwater"].copy()
    area = w.cell_area
    ow_area = (w.lake_area + w.reservoir_area) + (area / 100.0) * (w.loc_lake + w.loc_wetland + w.glo_wetland)
    surf = np.zeros(w.n)
    recharge = np.zeros(w.n)
    open_water = np.zeros(w.n)
    laf_pos = np.ones(w.n, bool)
    for d in range(n_days):
        laf = o.get_state()["laf"].copy()
        o.run(np.ascontiguousarray(f[d:d + 1]), d + 1)
        out = o.day_outputs()
        laf_pos &= laf > 0
        surf += out["surface_runoff"] * area / 1e6 * laf / 100.0
        recharge += out["gw_recharge"] * area * (laf / 100.0) / 1e6
        open_water += (out["prec"] - out["open_water_pet"]) * cfa / 1e6 * ow_area
    pot = o.pot_cell_runoff()
    expect = surf + recharge - (o.get_state()["groundwater"] - gw0) + open_water
    scale = np.abs(surf) + np.abs(recharge) + np.abs(open_water)
    swb = ow_area > 0
    kinds = {"humid": (w.aindex == 0) & (w.ldd >= 0), "inland sink": w.ldd < 0,
             "arid without surface water": (w.aindex == 1) & (w.ldd >= 0) & ~swb}
    for kind, sel in kinds.items():
        sel = sel & laf_pos
        assert sel.sum() >= 10, kind
        err = np.abs(pot[sel] - expect[sel]) / scale[sel]
        assert err.max() < 1e-12, (kind, float(err.max()))
    # the open-water term matters in the humid cells, the groundwater term in the arid ones
    assert np.median(np.abs(open_water[kinds["humid"]]) / scale[kinds["humid"]]) > 0.1
    sel = kinds["arid without surface water"]
    assert np.median(np.abs(recharge[sel]) / scale[sel]) > 0.01


def test_pot_cell_runoff_is_the_generated_volume(wg):
    check_pot_cell_runoff(wg)


# ---------------------------------------------------------------------------
# planted mistakes: every check above fails on the mutant it is written against
# ---------------------------------------------------------------------------

MUTANTS = [
    ("lake exponent 1.5 -> 2.5", [("#define LAKE_OUTFLOW_EXP 1.5", "#define LAKE_OUTFLOW_EXP 2.5")],
     check_local_lake),
    ("wetland exponent 2.5 -> 1.5", [("#define WETL_OUTFLOW_EXP 2.5", "#define WETL_OUTFLOW_EXP 1.5")],
     check_local_wetland),
    ("local lake reduction without 2 Smax",
     [("m->red_ll[i] = 1. - POW(fabs(m->loc_lake[i] - maxStorage) / (2. * maxStorage)",
       "m->red_ll[i] = 1. - POW(fabs(m->loc_lake[i] - maxStorage) / (maxStorage)")], check_local_lake),
    ("arid lake recharge dropped",
     [("gwr_loclak = KSWBGW * rf * st->loc_lake[i] / 100. / (cf / 100.);", "gwr_loclak = 0.;")],
     lambda wg, v: check_local_lake(wg, v, arid=True)),
    ("local lake evaporation without reduction factor",
     [("evap = ((1.0 - cfa) * ow_prec * rf) + (cfa * (ow_pet * rf));\n    if (evap < 0.) evap = 0.;\n    totalInflow = inflow + (ow_prec * rf) * (cellArea",
       "evap = ((1.0 - cfa) * ow_prec * rf) + (cfa * (ow_pet));\n    if (evap < 0.) evap = 0.;\n    totalInflow = inflow + (ow_prec * rf) * (cellArea")],
     check_local_lake),
    ("global wetland reduction over 2 Smax",
     [("m->red_gw[i] = 1. - POW(fabs(m->glo_wetl[i] - maxStorage) / maxStorage",
       "m->red_gw[i] = 1. - POW(fabs(m->glo_wetl[i] - maxStorage) / (2. * maxStorage)")], check_global_wetland),
    ("global lake inflow term halved",
     [("m->glo_lake[i] = prev * EXP(-1. * k_sw) + (1. / k_sw) * (totalInflow - use)",
       "m->glo_lake[i] = prev * EXP(-1. * k_sw) + (0.5 / k_sw) * (totalInflow - use)")], check_global_lake),
    ("river evaporation added", [("(1. / K) * (riverInflow - evapRemUse)", "(1. / K) * (riverInflow + evapRemUse)")],
     check_river),
    ("river hydraulic radius from the rectangle",
     [("double wp = bottom + 2.0 * depth * sqrt(5.0);", "double wp = bottom + 2.0 * depth;")], check_river),
    ("release coefficient over 0.8 cap",
     [("m->k_release[i] = m->res[i] / (maxStorage * 0.85);", "m->k_release[i] = m->res[i] / (maxStorage * 0.8);")],
     check_reservoir),
    ("10 % release rule dropped",
     [("outflow = 0.1 * release * (24. * 3600.) / 1000000000.;", "outflow = release * (24. * 3600.) / 1000000000.;")],
     check_reservoir),
    ("capacity-ratio blend 2c^2", [("((4. * c_ratio * c_ratio) * m->k_release[i] * prov_rel)",
                                    "((2. * c_ratio * c_ratio) * m->k_release[i] * prov_rel)")], check_reservoir),
    ("reservoir reduction exponent", [("#define EVAPO_RED_EXP_RES 2.81383", "#define EVAPO_RED_EXP_RES 3.32193")],
     check_reservoir),
    ("irrigation demand without the downstream cells",
     [("        m->mean_demand[i] += st->mean_nus[downstreamCell - 1] * coeff;\n", "\n")],
     check_irrigation_reservoir),
    ("irrigation demand walk testing the downstream cell itself",
     [("(downstreamCell < n ? st->reservoir_area[downstreamCell] : 0.) <= 0)",
       "st->reservoir_area[downstreamCell - 1] <= 0)")], check_irrigation_reservoir),
    ("irrigation release Q - D always",
     [("      if (mean_demand >= 0.5 * qmean)\n        prov_rel", "      if (0)\n        prov_rel")],
     check_irrigation_reservoir),
    ("local lake outflow not subtracted", [("    m->loc_lake[i] -= outflow;\n", "\n")], check_conservation),
    ("global wetland spill lost", [("      outflow += (m->glo_wetl[i] - maxStorage);\n", "\n")], check_conservation),
    ("reservoir negative storage not charged to outflow",
     [("    if (m->res[i] < 0.) {\n      outflow += m->res[i];", "    if (m->res[i] < 0.) {\n      outflow += 0.;")],
     check_conservation),
    ("river fraction change in fractions, not percent",
     [("(change_pct + (m->river_frac_change[i] * 100.))", "(change_pct + (m->river_frac_change[i]))")],
     check_land_area_identity),
    ("evergreen rule dropped", [("      if (lct <= 2) *status = 0;\n", "\n")], check_lai),
    ("full-cover reset one day short", [("        *days = 30 + init_days;", "        *days = 29 + init_days;")],
     check_lai),
    ("potential runoff of humid cells without the direct runoff",
     [("/* routing.cpp:2045-2053 */\n    m->pot_acc[i] += localRunoff + localRunoffIntoRiver +",
       "/* routing.cpp:2045-2053 */\n    m->pot_acc[i] += localRunoff +")], check_pot_cell_runoff),
    ("potential runoff of arid cells without the groundwater runoff",
     [("/* routing.cpp:3372-3379 */\n    m->pot_acc[i] += localRunoff + localRunoffIntoRiver + localGWRunoffIntoRiver +",
       "/* routing.cpp:3372-3379 */\n    m->pot_acc[i] += localRunoff + localRunoffIntoRiver +")], check_pot_cell_runoff),
    ("open-water balance without CFA",
     [("(m->ow_prec[i] - m->ow_pet[i]) * m->cfa[i];", "(m->ow_prec[i] - m->ow_pet[i]);")], check_pot_cell_runoff),
    ("threshold elevation ignored",
     [("          temp_elev = dailyTempC - ((thresh_elev - elev[0]) * T_GRADNT);", "          ;")],
     check_threshold_elevation),
    # water use (the checks live with the other water-use pins)
    ("return-flow factor without the groundwater fraction",
     [("      factor = 1 - (1 - frgi) * (1 - eff);", "      factor = 1 - frgi * (1 - eff);")],
     wu.check_net_abstraction_gw),
    ("reduced return flow not given back",
     [("    returnflowChange = (reintroducedRatio * USE(WG_USE_REDUCED_RETURN));",
       "    returnflowChange = 0.;")], wu.check_net_abstraction_gw),
    ("year end adds last year's unsatisfied use",
     [("      USE(WG_USE_TOTAL_UNSAT) -= USE(WG_USE_UNSAT_PREV_YEAR);",
       "      USE(WG_USE_TOTAL_UNSAT) += USE(WG_USE_UNSAT_PREV_YEAR);")], wu.check_year_end),
    ("reservoir use added, not taken",
     [("            m->res[i] -= remainingUseRes;\n            remainingUseRes = 0.;\n          } else {",
       "            m->res[i] += remainingUseRes;\n            remainingUseRes = 0.;\n          } else {")],
     wu.check_reservoir_use),
    ("river not asked for the use",
     [("  double evapRemUse = remainingUse + riverEvapo;", "  double evapRemUse = 0. + riverEvapo;")],
     wu.check_exhausting_use),
    ("local lake drawn to 0, not -Smax",
     [("        remainingUse -= (maxStorage + m->loc_lake[i]);\n        m->loc_lake[i] = (-1.) * maxStorage;",
       "        remainingUse -= (maxStorage + m->loc_lake[i]);\n        m->loc_lake[i] = 0.;")],
     wu.check_exhausting_use),
    ("use served before the river's evaporation",
     [("      remainingUse -= remainingUse * evapRemUseMax / evapRemUse;",
       "      remainingUse -= (evapRemUseMax < remainingUse ? evapRemUseMax : remainingUse);")],
     wu.check_river_share),
    ("global lake without its use",
     [("    use = evap * ((double)st->lake_area[i] / 1000000.) + gwr_glolak * cellArea * (cf / 100.) / 1000000. +\n          remainingUseGloLake;",
       "    use = evap * ((double)st->lake_area[i] / 1000000.) + gwr_glolak * cellArea * (cf / 100.) / 1000000. +\n          0.;")],
     wu.check_global_lake_use),
]


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    from concurrent.futures import ThreadPoolExecutor
    paths = [str(d / f"libwg_mutant{i}.so") for i in range(len(MUTANTS))]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        list(ex.map(lambda a: wg_oracle.build_variant(a[0][1], a[1]), zip(MUTANTS, paths)))
    return paths


@pytest.mark.parametrize("i", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_check_fails_on_planted_mistake(wg, mutant_libs, i):
    name, _repl, check = MUTANTS[i]
    with pytest.raises(AssertionError):
        check(wg, mutant_libs[i])
