"""ctypes mirror of include/watergap2_amd.h and watergap2_amd/csrc/wg_world.h.

Struct layouts follow the C declarations field by field; the library is the
single in-tree shared object ``watergap2_amd/libwatergap2_amd.so``.
"""
import ctypes as C
import os

import numpy as np

NSUB = 100
NELEV = 101
NLCT = 18
# water-use state fields, [WG_USE_NFIELDS][N] (include/watergap2_amd.h)
USE_FIELDS = ["total_unsat", "prev_total_unsat", "daily_remaining", "unsat_irrig", "unsat_other",
              "reduced_return", "withdrawal_irrig", "consumptive_irrig", "unsat_prev_year",
              "unsat_irrig_prev_year", "unsat_other_prev_year", "reduced_return_prev_year"]
WG_USE_NFIELDS = len(USE_FIELDS)

PARAMS = [
    "P_GAMRUN_C", "P_CFA", "P_CFS", "M_ROOT_D", "M_RIVRGH_C", "P_LAK_D", "P_WET_D",
    "P_SWOUTF_C", "M_EVAREDEX", "M_NETRAD", "P_PTC_HUM", "P_PTC_ARI", "P_PET_MXDY", "P_MCWH",
    "M_LAI", "P_T_SNOWFZ", "P_T_SNOWMT", "M_DEGDAY_F", "P_T_GRADNT", "M_GW_F", "M_RG_MAX",
    "P_PCRITGWA", "P_GWOUTF_C", "M_NETABSSW", "M_NETABSGW", "M_PREC",
]
NPARAM = len(PARAMS)
PARAM_INDEX = {name: i for i, name in enumerate(PARAMS)}

ERRORS = {0: "ok", -1: "invalid argument", -2: "cyclic or inconsistent flow graph",
          -3: "HIP runtime error", -4: "out of memory", -5: "invalid input value",
          -6: "call out of sequence"}

P_d = C.POINTER(C.c_double)
P_f = C.POINTER(C.c_float)
P_i32 = C.POINTER(C.c_int32)
P_i16 = C.POINTER(C.c_int16)
P_i8 = C.POINTER(C.c_int8)
P_u64 = C.POINTER(C.c_uint64)


class LctTables(C.Structure):
    _fields_ = [
        ("rooting_depth", C.c_float * NLCT),
        ("albedo", C.c_double * NLCT),
        ("albedo_snow", C.c_double * NLCT),
        ("ddf", C.c_double * NLCT),
        ("emissivity", C.c_double * NLCT),
        ("lai", C.c_float * NLCT),
        ("decid_frac", C.c_float * NLCT),
        ("evergreen_red", C.c_float * NLCT),
        ("initial_days", C.c_int16 * NLCT),
        ("kc_min", C.c_double * NLCT),
        ("kc_max", C.c_double * NLCT),
    ]


STATIC_FIELDS = [
    ("row", P_i16), ("cell_area", P_d), ("contfreq", P_d),
    ("aindex", P_i16), ("ldd", P_i8), ("land_cover", P_i8), ("built_up", P_f), ("smax", P_f),
    ("gw_factor", P_f), ("rgmax", P_i16), ("texture", P_i8), ("elevation", P_i16), ("params", P_d),
    ("glo_lake", P_d), ("loc_lake", P_d), ("glo_wetland", P_d), ("loc_wetland", P_d),
    ("lake_area", P_d), ("reservoir_area", P_d), ("glo_res", P_d), ("stor_cap", P_d),
    ("mean_outflow", P_d), ("res_type", P_i8), ("start_month", P_i8), ("river_length", P_f),
    ("river_slope", P_f), ("roughness", P_f), ("bankfull_flow", P_f), ("downstream", P_i32),
    ("rout_order", P_i32), ("lct", C.POINTER(LctTables)), ("mean_nus", P_d), ("alloc_coeff", P_d),
]


class StaticFields(C.Structure):
    _fields_ = STATIC_FIELDS


class OptionsProfile(C.Structure):
    _fields_ = [(name, C.c_int32) for name in (
        "time_series", "cloud", "intercept", "calc_albedo", "pet_opt", "use_kc", "river_velo_opt",
        "subtract_use", "res_opt", "statcorr_opt", "arid_area_opt", "fractional_routing_opt",
        "river_evapo_opt", "aggr_nus_glolakres_opt")]


# (name, ctype pointer, numpy dtype, channels)
STATE_FIELDS = [
    ("canopy", P_d, np.float64, 1), ("soil", P_d, np.float64, 1), ("snow", P_d, np.float64, 1),
    ("snow_elev", P_d, np.float64, NELEV), ("storage_transfer", P_d, np.float64, 1),
    ("lai_days", P_i32, np.int32, 1), ("lai_status", P_i32, np.int32, 1),
    ("lai_precsum", P_d, np.float64, 1),
    ("groundwater", P_d, np.float64, 1), ("loc_lake", P_d, np.float64, 1),
    ("loc_wetl", P_d, np.float64, 1), ("glo_lake", P_d, np.float64, 1),
    ("glo_wetl", P_d, np.float64, 1), ("reservoir", P_d, np.float64, 1),
    ("river", P_d, np.float64, 1),
    ("red_loc_lake", P_d, np.float64, 1), ("red_loc_wetl", P_d, np.float64, 1),
    ("red_glo_lake", P_d, np.float64, 1), ("red_glo_wetl", P_d, np.float64, 1),
    ("red_res", P_d, np.float64, 1), ("red_river", P_d, np.float64, 1),
    ("river_frac_next", P_d, np.float64, 1), ("fswb_frac_next", P_d, np.float64, 1),
    ("laf", P_d, np.float64, 1), ("laf_prev", P_d, np.float64, 1),
    ("k_release", P_d, np.float64, 1),
]


class StateStruct(C.Structure):
    _fields_ = [(f[0], f[1]) for f in STATE_FIELDS] + [("started", C.c_int32)]


DAY_OUTPUT_FIELDS = ["discharge", "surface_runoff", "gw_recharge", "open_water_pet",
                     "cell_aet_uncorr", "prec"]


class DayOutputs(C.Structure):
    _fields_ = [(f, P_d) for f in DAY_OUTPUT_FIELDS]


class WorldSpec(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("nrows", C.c_int32), ("ncols", C.c_int32), ("row0", C.c_int32),
        ("col0", C.c_int32), ("n_land", C.c_int32), ("n_outlets", C.c_int32),
        ("n_sinks", C.c_int32), ("seed", C.c_uint64), ("water_frac_max", C.c_double),
        ("glolake_every", C.c_int32), ("res_every", C.c_int32), ("edge_cases", C.c_int32),
        ("relief_scale", C.c_double),
    ]


WORLD_ARRAYS = [
    # name, ctype pointer, dtype, channels
    ("row", P_i16, np.int16, 1), ("col", P_i16, np.int16, 1), ("cell_at", P_i32, np.int32, None),
    ("cell_area", P_d, np.float64, 1), ("contfreq", P_d, np.float64, 1),
    ("aindex", P_i16, np.int16, 1), ("ldd", P_i8, np.int8, 1), ("land_cover", P_i8, np.int8, 1),
    ("texture", P_i8, np.int8, 1), ("res_type", P_i8, np.int8, 1),
    ("start_month", P_i8, np.int8, 1), ("built_up", P_f, np.float32, 1),
    ("smax", P_f, np.float32, 1), ("gw_factor", P_f, np.float32, 1),
    ("rgmax", P_i16, np.int16, 1), ("elevation", P_i16, np.int16, NELEV),
    ("params", P_d, np.float64, "params"),
    ("glo_lake", P_d, np.float64, 1), ("loc_lake", P_d, np.float64, 1),
    ("glo_wetland", P_d, np.float64, 1), ("loc_wetland", P_d, np.float64, 1),
    ("lake_area", P_d, np.float64, 1), ("reservoir_area", P_d, np.float64, 1),
    ("glo_res", P_d, np.float64, 1), ("stor_cap", P_d, np.float64, 1),
    ("mean_outflow", P_d, np.float64, 1),
    ("river_length", P_f, np.float32, 1), ("river_slope", P_f, np.float32, 1),
    ("roughness", P_f, np.float32, 1), ("bankfull_flow", P_f, np.float32, 1),
    ("altitude", P_f, np.float32, 1), ("meander", P_f, np.float32, 1),
    ("inflow9", P_i32, np.int32, 9), ("downstream", P_i32, np.int32, 1),
    ("rout_order", P_i32, np.int32, 1), ("level", P_i32, np.int32, 1),
    ("basin", P_i32, np.int32, 1),
]


class WorldStruct(C.Structure):
    _fields_ = [("n", C.c_int32)] + [
        (f[0], f[1]) for f in WORLD_ARRAYS if f[0] not in ("basin",)
    ] + [
        ("basin", P_i32), ("n_basins", C.c_int32), ("basin_outlet", P_i32),
        ("max_level", C.c_int32), ("upstream_area", P_d), ("lct", LctTables),
        ("kind", C.c_int32), ("seed", C.c_uint64),
    ]


def ptr(arr, ctype_ptr):
    """Pointer to a C-contiguous numpy array (None stays NULL)."""
    if arr is None:
        return None
    assert arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(ctype_ptr)


def lib_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "libwatergap2_amd.so")


_LIB = None


def load_library():
    """Loads the in-tree library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path)
    _declare(lib)
    _LIB = lib
    return lib


class GammaSearch(C.Structure):
    """wg_gamma_search (include/watergap2_amd.h)."""
    _fields_ = [("gamma_low", C.c_float), ("gamma_high", C.c_float), ("sum_diff_old", C.c_float),
                ("call_counter", C.c_int32), ("gamma_cond", C.c_int32), ("calib_status", C.c_int32),
                ("cell_corr_factor_ind", C.c_int32), ("cell_corr_factor", C.c_float),
                ("gamma_upper_limit", C.c_float), ("gamma_lower_limit", C.c_float),
                ("last_nash_sutcliffe", C.c_float), ("last_sum_diff", C.c_float),
                ("measured_avg_adapt", C.c_float), ("sim_runoff_avg", C.c_float)]


class RoutingInputs(C.Structure):
    """wg_routing_inputs (include/watergap2_amd.h)."""
    _fields_ = [("n_cells", C.c_int32), ("flowdir", P_i16), ("ldd", P_i8), ("row", P_i16), ("col", P_i16),
                ("cell_at", P_i32), ("altitude", P_f), ("meander", P_f), ("reservoir_area", P_d),
                ("mean_outflow", P_d), ("mean_outflow_monthly", P_d)]


class RoutingFiles(C.Structure):
    """wg_routing_files (include/watergap2_amd.h)."""
    _fields_ = [("ldd", P_i8), ("inflow9", P_i32), ("outflow", P_i32), ("flow_acc", P_i16),
                ("basins", C.POINTER(C.c_uint16)), ("cells_to_outlet", C.POINTER(C.c_uint16)),
                ("cell_dist", P_f), ("river_slope", P_f), ("river_length", P_f), ("rout_order", P_i32),
                ("alloc_coeff", P_d), ("start_month", P_i8), ("n_basins", C.c_int32)]


def _declare(lib):
    i32 = C.c_int32
    P_u16 = C.POINTER(C.c_uint16)
    lib.wg_ldd_from_flowdir.argtypes = [i32, P_i16, P_i8]
    lib.wg_replace_flow_to_ocean.argtypes = [i32, P_i16, P_i16, P_i32, P_i8]
    lib.wg_flow_accumulation.argtypes = [i32, P_i32, P_i16]
    lib.wg_derive_basins.argtypes = [i32, P_i8, P_i32, P_u16, P_u16, P_i32]
    lib.wg_reindex_basins.argtypes = [i32, i32, P_u16, P_i32]
    lib.wg_neighbour_cells.argtypes = [i32, P_i16, P_i16, P_i32, P_i32]
    lib.wg_cell_distances.argtypes = [P_f]
    lib.wg_river_slope.argtypes = [i32, P_f, P_f, P_f, P_i32, P_i8, P_i16, P_f]
    lib.wg_river_length.argtypes = [i32, P_f, P_i16, P_i16, P_f, P_i32, P_f]
    lib.wg_alloc_coeff.argtypes = [i32, P_i32, P_d, P_d, P_d]
    lib.wg_reservoir_start_month.argtypes = [i32, P_d, P_d, P_i8]
    lib.wg_prepare_routing.argtypes = [C.POINTER(RoutingInputs), C.POINTER(RoutingFiles)]
    lib.wg_unf_type.argtypes = [C.c_char_p]
    lib.wg_unf_read.argtypes = [C.c_char_p, i32, C.c_int64, C.c_void_p]
    lib.wg_unf_write.argtypes = [C.c_char_p, i32, C.c_int64, C.c_void_p]
    lib.wg_unf_read_double.argtypes = [C.c_char_p, C.c_int64, P_d]
    lib.wg_unf_write_double.argtypes = [C.c_char_p, C.c_int64, P_d]
    lib.wg_read_climate_year.argtypes = [C.c_char_p, i32, i32, P_f]
    lib.wg_read_lct_tables.argtypes = [C.c_char_p, C.POINTER(LctTables)]
    lib.wg_cell_geometry.argtypes = [i32, P_d, P_i16, P_d, P_d, P_d, P_d, P_d, P_d, P_d]
    lib.wg_soil_capacity.argtypes = [i32, P_f, P_i8, P_d, P_f, P_f]
    lib.wg_gw_factor.argtypes = [i32, P_i8, P_i8, P_i8, P_i8, P_f, P_d, P_d, i32, P_f, P_i16]
    for fn in ("wg_write_wghm_state_txt", "wg_read_wghm_state_txt", "wg_write_snow_elevation_txt",
               "wg_read_snow_elevation_txt", "wg_write_additional_txt", "wg_read_additional_txt"):
        getattr(lib, fn).argtypes = [C.c_char_p, i32, P_d]
    lib.wg_state_to_files.argtypes = [i32, C.POINTER(StaticFields), C.POINTER(StateStruct), P_d, P_d, P_d]
    lib.wg_state_from_files.argtypes = [i32, C.POINTER(StaticFields), P_d, P_d, P_d, C.POINTER(StateStruct)]
    lib.wg_last_error.restype = C.c_char_p
    lib.wg_last_error.argtypes = [C.c_void_p]
    lib.wg_default_options.argtypes = [C.POINTER(OptionsProfile)]
    lib.wg_default_options.restype = None
    lib.wg_inflow_outflow.argtypes = [i32, P_i8, P_i16, P_i16, P_i32, P_i32, P_i32]
    lib.wg_rout_order.argtypes = [i32, P_i32, P_i32, P_i32, P_i32]
    lib.wg_create.argtypes = [i32, C.POINTER(StaticFields), C.POINTER(OptionsProfile), i32,
                              C.POINTER(C.c_void_p)]
    lib.wg_destroy.argtypes = [C.c_void_p]
    lib.wg_get_state.argtypes = [C.c_void_p, i32, C.POINTER(StateStruct)]
    lib.wg_set_state.argtypes = [C.c_void_p, i32, C.POINTER(StateStruct)]
    lib.wg_set_member_params.argtypes = [C.c_void_p, P_d, P_d, P_f]
    lib.wg_set_forcing_perturbation.argtypes = [C.c_void_p, P_u64, C.c_double, C.c_double]
    lib.wg_load_forcing.argtypes = [C.c_void_p, P_f, i32]
    lib.wg_set_stations.argtypes = [C.c_void_p, i32, P_i32]
    lib.wg_step_days.argtypes = [C.c_void_p, i32, i32, i32]
    lib.wg_step_days_profiled.argtypes = [C.c_void_p, i32, i32, i32]
    lib.wg_schedule_info.argtypes = [C.c_void_p, P_i32]
    lib.wg_partitions.argtypes = [C.c_void_p, i32, P_i32, P_i32]
    lib.wg_libm_eval.argtypes = [i32, i32, P_d, P_d, P_d]
    lib.wg_get_station_discharge.argtypes = [C.c_void_p, P_d, P_d, i32]
    lib.wg_reset_station_discharge.argtypes = [C.c_void_p]
    lib.wg_set_pot_cell_runoff.argtypes = [C.c_void_p, i32]
    lib.wg_get_pot_cell_runoff.argtypes = [C.c_void_p, i32, P_d]
    lib.wg_set_water_use.argtypes = [C.c_void_p, C.POINTER(StaticFields), i32, i32, P_d, P_d, P_d, P_d, P_d, i32]
    lib.wg_get_water_use_state.argtypes = [C.c_void_p, i32, P_d]
    lib.wg_set_water_use_state.argtypes = [C.c_void_p, i32, P_d]
    lib.wg_station_daily_window.argtypes = [C.c_void_p, P_i32, P_i32]
    lib.wg_get_day_outputs.argtypes = [C.c_void_p, i32, C.POINTER(DayOutputs)]
    lib.wg_synchronize.argtypes = [C.c_void_p]
    lib.wg_stream.argtypes = [C.c_void_p]
    lib.wg_stream.restype = C.c_void_p
    lib.wg_last_timing.argtypes = [C.c_void_p, P_d, C.POINTER(C.c_int64)]
    lib.wg_autotune.argtypes = [C.c_void_p, i32, P_i32, P_d]
    lib.wg_set_launch_config.argtypes = [C.c_void_p, i32, i32, i32, i32]
    lib.wg_gamma_search_init.argtypes = [C.POINTER(GammaSearch)]
    lib.wg_gamma_search_init.restype = None
    lib.wg_gamma_start.argtypes = [C.POINTER(GammaSearch), C.c_float]
    lib.wg_gamma_start.restype = C.c_float
    lib.wg_find_new_gamma.argtypes = [C.POINTER(GammaSearch), C.c_float, i32, P_f, P_f, P_f, P_f]
    lib.wg_find_new_gamma.restype = C.c_float
    lib.wg_station_corr_factor.argtypes = [C.c_void_p, C.c_float, i32, P_f, P_f]
    lib.wg_station_corr_factor.restype = C.c_float
    lib.wg_calib_basins.argtypes = [i32, P_i32, P_i32, i32, P_i32, P_i32, P_i16, P_i32]
    lib.wg_upstream_stations.argtypes = [i32, P_i32, i32, P_i32, i32, P_i32, P_i32, i32]
    lib.wg_cell_corr_factor_grid.argtypes = [i32, P_i16, i32, i32, P_f, C.c_float, C.c_float, P_f, P_d]
    lib.wg_annual_init.argtypes = [C.c_void_p, C.POINTER(StaticFields), P_i32, i32]
    lib.wg_reservoir_year.argtypes = [i32, i32, P_d, P_d, P_i32, P_i8, P_d, P_d]
    lib.wg_mlp_create.argtypes = [C.c_uint64, C.POINTER(C.c_void_p)]
    lib.wg_mlp_destroy.argtypes = [C.c_void_p]
    lib.wg_mlp_last_error.argtypes = [C.c_void_p]
    lib.wg_mlp_last_error.restype = C.c_char_p
    lib.wg_mlp_n_params.argtypes = []
    lib.wg_mlp_n_params.restype = C.c_int64
    lib.wg_mlp_get_params.argtypes = [C.c_void_p, P_f]
    lib.wg_mlp_set_params.argtypes = [C.c_void_p, P_f]
    lib.wg_mlp_forward.argtypes = [C.c_void_p, i32, P_f, P_f]
    lib.wg_mlp_forward_noalloc.argtypes = [C.c_void_p, i32, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.wg_mlp_gradient.argtypes = [C.c_void_p, i32, P_f, P_f, P_f, P_f]
    lib.wg_mlp_train.argtypes = [C.c_void_p, i32, P_f, P_f, i32, i32, C.c_float, C.c_uint64, P_f]
    lib.wg_world_create.argtypes = [C.POINTER(WorldSpec), C.POINTER(C.POINTER(WorldStruct))]
    lib.wg_world_destroy.argtypes = [C.POINTER(WorldStruct)]
    lib.wg_world_destroy.restype = None
    lib.wg_world_statics.argtypes = [C.POINTER(WorldStruct), C.POINTER(StaticFields)]
    lib.wg_world_statics.restype = None
    lib.wg_world_forcing.argtypes = [C.POINTER(WorldStruct), C.c_uint64, C.c_int64, i32, P_f]
    lib.wg_world_subset.argtypes = [C.POINTER(WorldStruct), i32, P_i32,
                                    C.POINTER(C.POINTER(WorldStruct))]


# Every symbol declared in include/watergap2_amd.h (checked by the CPU tests).
ABI_SYMBOLS = [
    "wg_default_options", "wg_inflow_outflow", "wg_rout_order",
    "wg_ldd_from_flowdir", "wg_replace_flow_to_ocean", "wg_flow_accumulation", "wg_derive_basins", "wg_reindex_basins", "wg_neighbour_cells",
    "wg_cell_distances", "wg_river_slope", "wg_river_length", "wg_alloc_coeff", "wg_reservoir_start_month",
    "wg_prepare_routing", "wg_unf_type", "wg_unf_read", "wg_unf_write", "wg_unf_read_double",
    "wg_unf_write_double", "wg_read_climate_year", "wg_read_lct_tables", "wg_cell_geometry", "wg_soil_capacity",
    "wg_gw_factor", "wg_write_wghm_state_txt", "wg_read_wghm_state_txt", "wg_write_snow_elevation_txt",
    "wg_read_snow_elevation_txt", "wg_write_additional_txt", "wg_read_additional_txt", "wg_state_to_files",
    "wg_state_from_files", "wg_create", "wg_destroy",
    "wg_last_error", "wg_get_state", "wg_set_state", "wg_set_member_params",
    "wg_set_forcing_perturbation", "wg_load_forcing", "wg_set_stations", "wg_step_days",
    "wg_get_station_discharge", "wg_station_daily_window", "wg_reset_station_discharge", "wg_get_day_outputs",
    "wg_synchronize", "wg_stream", "wg_step_days_profiled", "wg_last_timing", "wg_schedule_info", "wg_libm_eval",
    "wg_autotune", "wg_set_launch_config",
    "wg_gamma_search_init", "wg_gamma_start", "wg_find_new_gamma",
    "wg_mlp_create", "wg_mlp_destroy", "wg_mlp_last_error", "wg_mlp_n_params", "wg_mlp_get_params",
    "wg_station_corr_factor", "wg_calib_basins", "wg_upstream_stations", "wg_annual_init", "wg_reservoir_year",
    "wg_set_pot_cell_runoff", "wg_get_pot_cell_runoff", "wg_cell_corr_factor_grid", "wg_partitions",
    "wg_set_water_use", "wg_get_water_use_state", "wg_set_water_use_state",
    "wg_mlp_set_params", "wg_mlp_forward", "wg_mlp_forward_noalloc", "wg_mlp_gradient", "wg_mlp_train",
]
