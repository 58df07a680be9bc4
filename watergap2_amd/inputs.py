"""WaterGAP input sets: the reference's file formats and file names, so that a
model runs from a WaterGAP input directory as the reference's does.

* :func:`unf_read` / :func:`unf_write` -- the .UNF grid format
  (grid_io_adapters.h:43-129: raw big-endian arrays; .UNF0 float32, .UNF1 int8,
  .UNF2 int16, .UNF4 int32; double grids are stored as float32).
* :func:`read_climate_year` -- the four daily .365.UNF0 forcing files of a year
  (climateYear.cpp:38-54) in the layout of ``WaterBalanceModel.load_forcing``.
* :func:`read_lct_tables` -- LCT_22.DAT and LAI_22.DAT (daily.cpp:1925-1959,
  lai.cpp:110-149).
* :func:`read_parameter_json` -- the per-cell parameter file
  (calib_param.h:131-222) as the [26][N] parameter array.
* :func:`load_watergap_inputs` -- the static inputs of the day loop for the
  parity option profile (antNatOpt 0, resOpt 1): the grids routingClass::init,
  geoClass::init, landClass, soilWatCapClass and groundwaterFactorClass read,
  with their derivations (surface-water fractions, unit conversions, Smax, fg,
  Rgmax), as an object that :class:`WaterBalanceModel` accepts like a World.
* :func:`write_watergap_inputs` -- the inverse: writes a grid's statics as a
  WaterGAP input set (used to run the synthetic grids through the file path).

The file I/O and the derivations run in the library (wg_inputs.cpp); this module
holds the file names and the order of the steps.
"""
import ctypes as C
import json
import os

import numpy as np

from . import _abi
from ._abi import NELEV, NLCT, NPARAM, P_d, P_f, P_i8, P_i16, P_i32

_TYPES = {0: (np.float32, P_f), 1: (np.int8, P_i8), 2: (np.int16, P_i16), 4: (np.int32, P_i32)}

# calib_param.h:131-222: JSON key of each parameter, in eCalibParam order
PARAMETER_KEYS = [
    "gammaHBV_runoff_coeff", "CFA_cellCorrFactor", "CFS_statCorrFactor", "root_depth_multiplier",
    "river_roughness_coeff_mult", "lake_depth", "wetland_depth", "surfacewater_outflow_coefficient",
    "evapo_red_fact_exp_mult", "net_radiation_mult", "PT_coeff_humid", "PT_coeff_arid", "max_daily_PET", "mcwh",
    "LAI_mult", "snow_freeze_temp", "snow_melt_temp", "degree_day_factor_mult", "temperature_gradient",
    "gw_factor_mult", "rg_max_mult", "pcrit_aridgw", "groundwater_outflow_coeff",
    "net_abstraction_surfacewater_mult", "net_abstraction_groundwater_mult", "precip_mult",
]
assert len(PARAMETER_KEYS) == NPARAM


def _lib():
    return _abi.load_library()


def _check(rc, what=""):
    if rc != 0:
        from . import WaterGapError
        raise WaterGapError(rc, what)


def unf_type(path):
    t = _lib().wg_unf_type(os.fsencode(path))
    _check(t if t < 0 else 0, f"not a .UNF file name: {path}")
    return t


def unf_read(path, count):
    """The first `count` elements of a .UNF file, typed by its extension."""
    t = unf_type(path)
    dt, ct = _TYPES[t]
    out = np.empty(count, dt)
    _check(_lib().wg_unf_read(os.fsencode(path), t, count, _abi.ptr(out, ct)), path)
    return out


def unf_read_double(path, count):
    """A double grid stored in a .UNF0 file (float32 on disk)."""
    out = np.empty(count, np.float64)
    _check(_lib().wg_unf_read_double(os.fsencode(path), count, _abi.ptr(out, P_d)), path)
    return out


def unf_write(path, arr):
    """Writes arr in the element type of the file's extension (doubles to .UNF0
    are rounded to float32, as the reference stores double grids)."""
    t = unf_type(path)
    dt, ct = _TYPES[t]
    arr = np.asarray(arr)
    if t == 0 and arr.dtype == np.float64:
        a = np.ascontiguousarray(arr.ravel())
        _check(_lib().wg_unf_write_double(os.fsencode(path), a.size, _abi.ptr(a, P_d)), path)
        return
    a = np.ascontiguousarray(arr.ravel(), dtype=dt)
    _check(_lib().wg_unf_write(os.fsencode(path), t, a.size, _abi.ptr(a, ct)), path)


CLIMATE_FILES = ["G_TEMP_H08_int_{}.365.UNF0", "G_GPCC_H08day_V20110128_{}.365.UNF0", "G_SSRD_H08_int_{}.365.UNF0",
                 "G_SLRD_H08_int_{}.365.UNF0"]


def read_climate_year(climate_dir, year, n_cells):
    """[365, 4, N] float32 forcing (T degC, P mm/d, SW, LW W/m2) of one year."""
    out = np.empty((365, 4, n_cells), np.float32)
    _check(_lib().wg_read_climate_year(os.fsencode(climate_dir), int(year), n_cells, _abi.ptr(out, P_f)),
           f"climate files of {year} in {climate_dir}")
    return out


def write_climate_year(climate_dir, year, forcing):
    """Writes [365, 4, N] forcing as the reference's four [N][365] files."""
    for v, name in enumerate(CLIMATE_FILES):
        unf_write(os.path.join(climate_dir, name.format(year)),
                  np.ascontiguousarray(forcing[:, v, :].T).astype(np.float32))


def read_lct_tables(input_dir):
    t = _abi.LctTables()
    _check(_lib().wg_read_lct_tables(os.fsencode(input_dir), C.byref(t)), f"LCT_22.DAT / LAI_22.DAT in {input_dir}")
    return t


def write_lct_tables(input_dir, t):
    """LCT_22.DAT and LAI_22.DAT in the reference's layout (repr() keeps every
    value's bits)."""
    f32 = lambda v: repr(float(np.float32(v)))  # noqa: E731
    with open(os.path.join(input_dir, "LCT_22.DAT"), "w") as f:
        f.write("# land cover class, rooting depth, albedo, snow albedo, degree-day factor, emissivity\n")
        for i in range(NLCT):
            f.write(f"{i + 1} {f32(t.rooting_depth[i])} {t.albedo[i]!r} {t.albedo_snow[i]!r} {t.ddf[i]!r} "
                    f"{t.emissivity[i]!r}\n")
    with open(os.path.join(input_dir, "LAI_22.DAT"), "w") as f:
        f.write("# land cover class, LAI, deciduous fraction, evergreen LAI reduction, initial days, kc min, "
                "kc max\n")
        for i in range(NLCT):
            f.write(f"{i + 1} {f32(t.lai[i])} {f32(t.decid_frac[i])} {f32(t.evergreen_red[i])} "
                    f"{t.initial_days[i]} {t.kc_min[i]!r} {t.kc_max[i]!r}\n")


def read_parameter_json(path):
    """[26][N] parameters from the reference's JSON parameter file (one array of N
    numbers per key, calib_param.h:131-222)."""
    with open(path) as f:
        js = json.load(f)
    return np.array([np.asarray(js[k], dtype=np.float64) for k in PARAMETER_KEYS])


def write_parameter_json(path, params):
    js = {"ng_param": int(params.shape[1])}
    for k, row in zip(PARAMETER_KEYS, params):
        js[k] = [float(v) for v in row]
    with open(path, "w") as f:
        json.dump(js, f)


class InputWorld:
    """Static inputs read from a WaterGAP input set; accepted by
    :class:`WaterBalanceModel` like a :class:`World`."""

    def static_fields(self):
        from . import World
        return World.static_fields(self)


def load_watergap_inputs(input_dir, routing_dir, params, res_year=None, nus_file=None, time_series=1):
    """Reads the static inputs of the day loop (parity option profile: anthropogenic
    run with the reservoir algorithm, antNatOpt 0, resOpt 1):

    * geometry (geo.cpp:7-44): GR, GC, GCRC, GAREA (by row) and GCONTFREQ, raised to
      the sum of G_LOCLAK, G_GLOLAK, G_LOCRES and G_RES/G_RES_FRAC;
    * surface water (routing.cpp:252-343, 424-425): G_GLOLAK, G_LOCLAK + G_LOCRES,
      G_GLOWET, G_LOCWET, G_LAKAREA, G_RESAREA (negative values 0), the reservoir
      fraction G_RES/G_RES_<res_year> (or G_RES_FRAC), G_RES_TYPE,
      G_STORAGE_CAPACITY, G_MEAN_OUTFLOW converted from km3/month to m3/s, the
      start months and allocation coefficients of the routing directory and the
      mean net use ``nus_file`` (e.g. G_NUs_1971_2000.UNF0);
    * rivers (routing.cpp:476-490): G_RIVER_LENGTH, G_RIVERSLOPE, G_ROUGHNESS,
      G_BANKFULL; routing order and outflow cells (G_ROUT_ORDER, G_OUTFLC);
    * land (daily.cpp, land.cpp): G_ARID_HUMID, G_LDD_2, G_LANDCOVER, GBUILTUP,
      G_ELEV_RANGE.101, LCT_22.DAT / LAI_22.DAT;
    * Smax (s_max.cpp:43-74) from G_TAWC, fg and Rgmax (gw_frac.cpp:31-258) from
      G_SLOPE_CLASS, G_TEXTURE, G_PERMAGLAC, G_AQ_FACTOR, G_GW_FACTOR_CORR.

    params is the [26][N] parameter array (see read_parameter_json)."""
    I = lambda name: os.path.join(input_dir, name)  # noqa: E731
    R = lambda name: os.path.join(routing_dir, name)  # noqa: E731
    w = InputWorld()
    row = unf_read(I("GR.UNF2"), _count(I("GR.UNF2"), 2))
    n = len(row)
    w.n = n
    w.row = row
    w.col = unf_read(I("GC.UNF2"), n)
    gcrc = unf_read(I("GCRC.UNF4"), 720 * 360).reshape(720, 360)  # GCRC(col, row): [720][360]
    w.cell_at = np.ascontiguousarray(gcrc.T).ravel()
    area = unf_read_double(I("GAREA.UNF0"), 360)
    dd = lambda name: unf_read_double(name, n)  # noqa: E731
    contfreq_in = dd(I("GCONTFREQ.UNF0"))
    loc_lake = np.maximum(dd(I("G_LOCLAK.UNF0")), 0.0)
    glo_lake = np.maximum(dd(I("G_GLOLAK.UNF0")), 0.0)
    loc_res = dd(I("G_LOCRES.UNF0"))
    res_frac = dd(I("G_RES/G_RES_FRAC.UNF0"))
    w.cell_area = np.empty(n)
    w.contfreq = np.empty(n)
    # geo.cpp adds the fractions as read (before routing's clamps)
    g_loc, g_glo = dd(I("G_LOCLAK.UNF0")), dd(I("G_GLOLAK.UNF0"))
    _check(_lib().wg_cell_geometry(n, _abi.ptr(area, P_d), _abi.ptr(row, P_i16), _abi.ptr(contfreq_in, P_d),
                                   _abi.ptr(g_loc, P_d), _abi.ptr(g_glo, P_d), _abi.ptr(loc_res, P_d),
                                   _abi.ptr(res_frac, P_d), _abi.ptr(w.cell_area, P_d), _abi.ptr(w.contfreq, P_d)))
    w.glo_lake = glo_lake
    w.loc_lake = loc_lake + loc_res  # routing.cpp:327 G_loc_lake += G_loc_res
    w.glo_wetland = np.maximum(dd(I("G_GLOWET.UNF0")), 0.0)
    w.loc_wetland = np.maximum(dd(I("G_LOCWET.UNF0")), 0.0)
    w.lake_area = np.maximum(dd(I("G_LAKAREA.UNF0")), 0.0)
    w.reservoir_area = np.maximum(dd(I("G_RESAREA.UNF0")), 0.0)
    w.glo_res = dd(I(f"G_RES/G_RES_{res_year}.UNF0")) if res_year is not None else res_frac
    stor = dd(I("G_STORAGE_CAPACITY.UNF0"))
    w.stor_cap = np.where(stor < 0.0, 0.0, stor)
    # routing.cpp:424-425: km3/month -> km3/yr -> m3/s
    w.mean_outflow = dd(I("G_MEAN_OUTFLOW.UNF0")) * 12.0 * 1000000000.0 / 31536000.0
    w.res_type = unf_read(I("G_RES_TYPE.UNF1"), n)
    w.start_month = unf_read(R("G_START_MONTH.UNF1"), n)
    w.mean_nus = dd(I(nus_file)) if nus_file else None
    w.alloc_coeff = unf_read_double(R("G_ALLOC_COEFF.5.UNF0"), n * 5) if os.path.exists(
        R("G_ALLOC_COEFF.5.UNF0")) else None
    w.river_length = unf_read(R("G_RIVER_LENGTH.UNF0"), n)
    w.river_slope = unf_read(R("G_RIVERSLOPE.UNF0"), n)
    w.roughness = unf_read(I("G_ROUGHNESS.UNF0"), n)
    w.bankfull_flow = unf_read(I("G_BANKFULL.UNF0"), n)
    w.downstream = unf_read(R("G_OUTFLC.UNF4"), n)
    w.rout_order = unf_read(R("G_ROUT_ORDER.UNF4"), n)
    w.aindex = unf_read(I("G_ARID_HUMID.UNF2"), n)
    w.ldd = unf_read(R("G_LDD_2.UNF1"), n)
    w.land_cover = unf_read(I("G_LANDCOVER.UNF1"), n)
    w.built_up = unf_read(I("GBUILTUP.UNF0"), n)
    w.elevation = unf_read(I("G_ELEV_RANGE.101.UNF2"), n * NELEV).reshape(n, NELEV)
    w.lct = read_lct_tables(input_dir)
    w.params = np.ascontiguousarray(params, dtype=np.float64)
    assert w.params.shape == (NPARAM, n)
    # Smax and the groundwater statics
    tawc = unf_read(I("G_TAWC.UNF0"), n)
    w.smax = np.empty(n, np.float32)
    rd = np.array([w.lct.rooting_depth[i] for i in range(NLCT)], np.float32)
    m_root = np.ascontiguousarray(w.params[3])
    _check(_lib().wg_soil_capacity(n, _abi.ptr(tawc, P_f), _abi.ptr(w.land_cover, P_i8), _abi.ptr(m_root, P_d),
                                   _abi.ptr(rd, P_f), _abi.ptr(w.smax, P_f)))
    w.texture = unf_read(I("G_TEXTURE.UNF1"), n)
    slope_class = unf_read(I("G_SLOPE_CLASS.UNF1"), n)
    permaglac = unf_read(I("G_PERMAGLAC.UNF1"), n)
    aq = unf_read(I("G_AQ_FACTOR.UNF1"), n)
    corr = unf_read(I("G_GW_FACTOR_CORR.UNF0"), n) if os.path.exists(I("G_GW_FACTOR_CORR.UNF0")) else None
    w.gw_factor = np.empty(n, np.float32)
    w.rgmax = np.empty(n, np.int16)
    m_gw, m_rg = np.ascontiguousarray(w.params[19]), np.ascontiguousarray(w.params[20])
    _check(_lib().wg_gw_factor(n, _abi.ptr(slope_class, P_i8), _abi.ptr(w.texture, P_i8), _abi.ptr(permaglac, P_i8),
                               _abi.ptr(aq, P_i8), _abi.ptr(corr, P_f), _abi.ptr(m_gw, P_d), _abi.ptr(m_rg, P_d),
                               int(time_series), _abi.ptr(w.gw_factor, P_f), _abi.ptr(w.rgmax, P_i16)))
    return w


def _count(path, size):
    return os.path.getsize(path) // size


def write_watergap_inputs(world, input_dir, routing_dir, tawc, slope_class, texture, permaglac, aq_factor,
                          gw_factor_corr=None, nus_file=None):
    """Writes a grid's statics as a WaterGAP input set.  Smax, fg and Rgmax are not
    files of the set: the reference derives them from the TAWC, slope class,
    texture, permafrost/glacier and aquifer grids given here.  The surface-water
    fractions go out with the local reservoirs' share 0 and the reservoir
    fraction as G_RES/G_RES_FRAC."""
    os.makedirs(os.path.join(input_dir, "G_RES"), exist_ok=True)
    os.makedirs(routing_dir, exist_ok=True)
    I = lambda name: os.path.join(input_dir, name)  # noqa: E731
    R = lambda name: os.path.join(routing_dir, name)  # noqa: E731
    n = world.n
    unf_write(I("GR.UNF2"), world.row)
    unf_write(I("GC.UNF2"), world.col)
    unf_write(I("GCRC.UNF4"), np.ascontiguousarray(world.cell_at.reshape(360, 720).T))
    area = np.zeros(360)
    area[world.row - 1] = world.cell_area
    unf_write(I("GAREA.UNF0"), area)
    unf_write(I("GCONTFREQ.UNF0"), np.asarray(world.contfreq, np.float64))
    for name, v in (("G_GLOLAK.UNF0", world.glo_lake), ("G_LOCLAK.UNF0", world.loc_lake),
                    ("G_LOCRES.UNF0", np.zeros(n)), ("G_RES/G_RES_FRAC.UNF0", world.glo_res),
                    ("G_GLOWET.UNF0", world.glo_wetland), ("G_LOCWET.UNF0", world.loc_wetland),
                    ("G_LAKAREA.UNF0", world.lake_area), ("G_RESAREA.UNF0", world.reservoir_area),
                    ("G_STORAGE_CAPACITY.UNF0", world.stor_cap),
                    ("G_MEAN_OUTFLOW.UNF0", np.asarray(world.mean_outflow) * 31536000.0 / (12.0 * 1000000000.0))):
        unf_write(I(name), np.asarray(v, np.float64))
    unf_write(I("G_RES_TYPE.UNF1"), world.res_type)
    unf_write(R("G_START_MONTH.UNF1"), world.start_month)
    if getattr(world, "mean_nus", None) is not None and nus_file:
        unf_write(I(nus_file), np.asarray(world.mean_nus, np.float64))
    if getattr(world, "alloc_coeff", None) is not None:
        unf_write(R("G_ALLOC_COEFF.5.UNF0"), np.asarray(world.alloc_coeff, np.float64))
    unf_write(R("G_RIVER_LENGTH.UNF0"), world.river_length)
    unf_write(R("G_RIVERSLOPE.UNF0"), world.river_slope)
    unf_write(I("G_ROUGHNESS.UNF0"), world.roughness)
    unf_write(I("G_BANKFULL.UNF0"), world.bankfull_flow)
    unf_write(R("G_OUTFLC.UNF4"), world.downstream)
    unf_write(R("G_ROUT_ORDER.UNF4"), world.rout_order)
    unf_write(I("G_ARID_HUMID.UNF2"), world.aindex)
    unf_write(R("G_LDD_2.UNF1"), world.ldd)
    unf_write(I("G_LANDCOVER.UNF1"), world.land_cover)
    unf_write(I("GBUILTUP.UNF0"), world.built_up)
    unf_write(I("G_ELEV_RANGE.101.UNF2"), world.elevation)
    write_lct_tables(input_dir, world.lct)
    unf_write(I("G_TAWC.UNF0"), tawc)
    unf_write(I("G_SLOPE_CLASS.UNF1"), slope_class)
    unf_write(I("G_TEXTURE.UNF1"), texture)
    unf_write(I("G_PERMAGLAC.UNF1"), permaglac)
    unf_write(I("G_AQ_FACTOR.UNF1"), aq_factor)
    if gw_factor_corr is not None:
        unf_write(I("G_GW_FACTOR_CORR.UNF0"), gw_factor_corr)
