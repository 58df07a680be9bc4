"""WaterGAP 2 daily water balance and lateral routing on MI355X (gfx950).

Host-side Python mirror of the C ABI in ``include/watergap2_amd.h``.  The compute
runs in ``libwatergap2_amd.so`` (HIP kernels); this module only marshals arrays.

* :class:`World` -- a seeded synthetic 0.5-degree grid (statics, routing schedule,
  forcing); the reference ships no input data.
* :class:`WaterBalanceModel` -- one model handle: the reference's day loop
  (``dailyWaterBalanceClass::calcNewDay`` for all cells, ``routingClass::routing``,
  ``routingClass::updateLandAreaFrac``; integrateWGHM.cpp:755-798) for one or more
  members.
* :func:`rout_order`, :func:`inflow_outflow` -- the routing schedule
  (rout_prepare.cpp:338-439, 619-670, 834-886).

Errors are raised as :class:`WaterGapError` carrying the C error code.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import NELEV, NPARAM, NSUB, PARAM_INDEX, PARAMS, STATE_FIELDS, DAY_OUTPUT_FIELDS

__all__ = ["World", "WaterBalanceModel", "WaterGapError", "rout_order", "inflow_outflow",
           "regional_spec", "global_spec", "PARAMS", "PARAM_INDEX", "NSUB", "NELEV", "NPARAM"]


class WaterGapError(RuntimeError):
    def __init__(self, code, message=""):
        self.code = code
        text = _abi.ERRORS.get(code, "error")
        super().__init__(f"{text} ({code}){': ' + message if message else ''}")


def _check(rc, handle=None):
    if rc != 0:
        msg = ""
        if handle:
            raw = _abi.load_library().wg_last_error(handle)
            msg = raw.decode() if raw else ""
        raise WaterGapError(rc, msg)


def regional_spec(seed=0, nrows=40, ncols=50, row0=81, col0=361, water_frac_max=0.1,
                  glolake_every=100, res_every=200, edge_cases=False):
    """Regional lattice, all land, one outlet (the 2,000-cell 50 x 40 workload)."""
    return _abi.WorldSpec(kind=0, nrows=nrows, ncols=ncols, row0=row0, col0=col0, n_land=0,
                          n_outlets=0, n_sinks=0, seed=seed, water_frac_max=water_frac_max,
                          glolake_every=glolake_every, res_every=res_every,
                          edge_cases=1 if edge_cases else 0)


def global_spec(seed=1, n_land=67420, n_outlets=600, n_sinks=20, water_frac_max=0.1,
                glolake_every=100, res_every=200, edge_cases=False, relief_scale=0.0):
    """Global land mask (the 67,420-cell 0.5-degree workload)."""
    return _abi.WorldSpec(kind=1, nrows=0, ncols=0, row0=0, col0=0, n_land=n_land,
                          n_outlets=n_outlets, n_sinks=n_sinks, seed=seed,
                          water_frac_max=water_frac_max, glolake_every=glolake_every,
                          res_every=res_every, edge_cases=1 if edge_cases else 0,
                          relief_scale=relief_scale)


class World:
    """Seeded synthetic grid produced by the library's generator."""

    def __init__(self, spec=None, _ptr=None):
        lib = _abi.load_library()
        if _ptr is None:
            ptr = C.POINTER(_abi.WorldStruct)()
            _check(lib.wg_world_create(C.byref(spec), C.byref(ptr)))
        else:
            ptr = _ptr
        self._lib = lib
        self._ptr = ptr
        w = ptr.contents
        self.n = int(w.n)
        self.n_basins = int(w.n_basins)
        self.max_level = int(w.max_level)
        self.kind = int(w.kind)
        self.seed = int(w.seed)
        n = self.n
        for name, _ct, dt, ch in _abi.WORLD_ARRAYS:
            if ch is None:
                count = 360 * 720
            elif ch == "params":
                count = NPARAM * n
            else:
                count = n * ch
            raw = getattr(w, name)
            arr = np.ctypeslib.as_array(raw, shape=(count,)).astype(dt, copy=True)
            if ch == "params":
                arr = arr.reshape(NPARAM, n)
            elif ch not in (None, 1):
                arr = arr.reshape(n, ch)
            setattr(self, name, arr)
        self.basin_outlet = np.ctypeslib.as_array(w.basin_outlet, shape=(self.n_basins,)).copy()
        self.upstream_area = np.ctypeslib.as_array(w.upstream_area, shape=(n,)).copy()
        self.lct = _abi.LctTables()
        C.memmove(C.byref(self.lct), C.byref(w.lct), C.sizeof(_abi.LctTables))
        # mean net surface-water use (m3/yr) and the allocation coefficients of the
        # five downstream cells ([N][5]) behind the irrigation reservoirs' demand;
        # None = no demand
        self.mean_nus = None
        self.alloc_coeff = None

    def __del__(self):
        try:
            if self._ptr:
                self._lib.wg_world_destroy(self._ptr)
                self._ptr = None
        except Exception:
            pass

    def forcing(self, seed, day0, n_days):
        """Daily forcing [n_days, 4, N] float32 (T, P, SW, LW) for day indices day0.."""
        out = np.empty((n_days, 4, self.n), dtype=np.float32)
        _check(self._lib.wg_world_forcing(self._ptr, seed, day0, n_days, _abi.ptr(out, _abi.P_f)))
        return out

    def subset(self, cells):
        """Sub-grid of the given 0-based cells (e.g. one basin)."""
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        ptr = C.POINTER(_abi.WorldStruct)()
        _check(self._lib.wg_world_subset(self._ptr, len(cells), _abi.ptr(cells, _abi.P_i32),
                                         C.byref(ptr)))
        return World(_ptr=ptr)

    def static_fields(self):
        """wg_static_fields pointing at this world's arrays (keep the world alive)."""
        s = _abi.StaticFields()
        for name, ct in _abi.STATIC_FIELDS:
            if name == "lct":
                s.lct = C.pointer(self.lct)
            else:
                arr = getattr(self, name)
                if arr is not None and name in ("mean_nus", "alloc_coeff"):
                    arr = np.ascontiguousarray(arr, dtype=np.float64)
                    setattr(self, name, arr)
                setattr(s, name, _abi.ptr(arr, ct))
        return s

    def basin_cells(self, b):
        return np.nonzero(self.basin == b)[0].astype(np.int32)


def inflow_outflow(world):
    lib = _abi.load_library()
    n = world.n
    inflow9 = np.zeros((n, 9), dtype=np.int32)
    outflow = np.zeros(n, dtype=np.int32)
    _check(lib.wg_inflow_outflow(n, _abi.ptr(world.ldd, _abi.P_i8), _abi.ptr(world.row, _abi.P_i16),
                                 _abi.ptr(world.col, _abi.P_i16), _abi.ptr(world.cell_at, _abi.P_i32),
                                 _abi.ptr(inflow9, _abi.P_i32), _abi.ptr(outflow, _abi.P_i32)))
    return inflow9, outflow


def rout_order(inflow9, outflow):
    """Routing order (1-based) and Kahn level; raises WaterGapError on a cycle."""
    lib = _abi.load_library()
    inflow9 = np.ascontiguousarray(inflow9, dtype=np.int32)
    outflow = np.ascontiguousarray(outflow, dtype=np.int32)
    n = len(outflow)
    order = np.zeros(n, dtype=np.int32)
    level = np.zeros(n, dtype=np.int32)
    _check(lib.wg_rout_order(n, _abi.ptr(inflow9, _abi.P_i32), _abi.ptr(outflow, _abi.P_i32),
                             _abi.ptr(order, _abi.P_i32), _abi.ptr(level, _abi.P_i32)))
    return order, level


def new_state(n):
    """Dict of zeroed state arrays (reference layout)."""
    st = {}
    for name, _ct, dt, ch in STATE_FIELDS:
        st[name] = np.zeros((n, ch) if ch > 1 else n, dtype=dt)
    st["started"] = 0
    return st


def state_struct(st):
    s = _abi.StateStruct()
    for name, ct, _dt, _ch in STATE_FIELDS:
        arr = st.get(name)
        setattr(s, name, _abi.ptr(arr, ct) if arr is not None else None)
    s.started = int(st.get("started", 0))
    return s


class WaterBalanceModel:
    """GPU model handle over the C ABI (one HIP stream, n_members members)."""

    def __init__(self, world, n_members=1, options=None):
        self._lib = _abi.load_library()
        self.world = world
        self.n = world.n
        self.n_members = n_members
        self._static = world.static_fields()
        opt = _abi.OptionsProfile()
        self._lib.wg_default_options(C.byref(opt))
        if options:
            for k, v in options.items():
                setattr(opt, k, v)
        h = C.c_void_p()
        rc = self._lib.wg_create(self.n, C.byref(self._static), C.byref(opt), n_members, C.byref(h))
        if rc != 0:
            msg = ""
            if h:
                msg = (self._lib.wg_last_error(h) or b"").decode()
                self._lib.wg_destroy(h)
            raise WaterGapError(rc, msg)
        self._h = h
        self.n_stations = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.wg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self._h)

    def load_forcing(self, forcing):
        forcing = np.ascontiguousarray(forcing, dtype=np.float32)
        assert forcing.ndim == 3 and forcing.shape[1:] == (4, self.n)
        self._c(self._lib.wg_load_forcing(self._h, _abi.ptr(forcing, _abi.P_f), forcing.shape[0]))

    def annual_init(self, year_statics, outflow_cell=None, apply_changes=True):
        """routingClass::annualInit for a year after the first (wg_annual_init):
        year_statics is the world with this year's glo_res, reservoir_area and
        stor_cap (any object with the World's static arrays)."""
        st = World.static_fields(year_statics)
        oc = None if outflow_cell is None else np.ascontiguousarray(outflow_cell, dtype=np.int32)
        self._c(self._lib.wg_annual_init(self._h, C.byref(st), None if oc is None else _abi.ptr(oc, _abi.P_i32),
                                         1 if apply_changes else 0))

    def set_stations(self, cells):
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.n_stations = len(cells)
        self._c(self._lib.wg_set_stations(self._h, len(cells), _abi.ptr(cells, _abi.P_i32)))

    def set_member_params(self, gamma=None, cfa=None, smax=None):
        def prep(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            assert a.size == self.n * self.n_members
            return a
        g, c, s = prep(gamma, np.float64), prep(cfa, np.float64), prep(smax, np.float32)
        self._c(self._lib.wg_set_member_params(self._h, _abi.ptr(g, _abi.P_d), _abi.ptr(c, _abi.P_d),
                                               _abi.ptr(s, _abi.P_f)))

    def set_forcing_perturbation(self, seeds, prec_sigma=0.2, temp_sigma=1.0):
        if seeds is None:
            self._c(self._lib.wg_set_forcing_perturbation(self._h, None, 0.0, 0.0))
            return
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        assert len(seeds) == self.n_members
        self._c(self._lib.wg_set_forcing_perturbation(self._h, _abi.ptr(seeds, _abi.P_u64),
                                                      prec_sigma, temp_sigma))

    def step_days(self, forcing_day0, first_doy, n_days, profiled=False):
        fn = self._lib.wg_step_days_profiled if profiled else self._lib.wg_step_days
        self._c(fn(self._h, forcing_day0, first_doy, n_days))

    def autotune(self, n_days=5):
        """Picks the fastest launch configuration on n_days of the loaded forcing
        (state restored afterwards); returns (cell threads, chain threads, snow
        chunk, chain mode) and the best time per day in ms."""
        cfg = (C.c_int32 * 4)()
        ms = C.c_double()
        self._c(self._lib.wg_autotune(self._h, n_days, cfg, C.byref(ms)))
        return tuple(cfg), ms.value

    def set_launch_config(self, cell_threads=256, chain_threads=512, snow_chunk=5, chain_mode=3):
        self._c(self._lib.wg_set_launch_config(self._h, cell_threads, chain_threads, snow_chunk, chain_mode))

    def timing(self):
        ms = (C.c_double * 3)()
        cnt = (C.c_int64 * 3)()
        self._c(self._lib.wg_last_timing(self._h, ms, cnt))
        return list(ms), list(cnt)

    def schedule_info(self):
        info = (C.c_int32 * 4)()
        self._lib.wg_schedule_info.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        self._c(self._lib.wg_schedule_info(self._h, info))
        return {"groups": info[0], "max_depth": info[1], "level_slots": info[2],
                "level0_cells": info[3]}

    def partitions(self):
        """[(cells, deepest group levels)] of the day loop's partitions (wg_partitions)."""
        cells = np.zeros(2, np.int32)
        depth = np.zeros(2, np.int32)
        n = self._lib.wg_partitions(self._h, 2, _abi.ptr(cells, _abi.P_i32), _abi.ptr(depth, _abi.P_i32))
        if n < 0:
            self._c(n)
        return [(int(cells[k]), int(depth[k])) for k in range(min(n, 2))]

    def get_state(self, member=0):
        st = new_state(self.n)
        s = state_struct(st)
        self._c(self._lib.wg_get_state(self._h, member, C.byref(s)))
        st["started"] = int(s.started)
        return st

    def set_state(self, st, member=-1):
        full = new_state(self.n)
        for k in full:
            if k in st and k != "started":
                full[k] = np.ascontiguousarray(st[k], dtype=full[k].dtype).reshape(full[k].shape)
        full["started"] = st.get("started", 0)
        s = state_struct(full)
        self._c(self._lib.wg_set_state(self._h, member, C.byref(s)))

    def day_outputs(self, member=0):
        out = {k: np.zeros(self.n) for k in DAY_OUTPUT_FIELDS}
        s = _abi.DayOutputs(*[_abi.ptr(out[k], _abi.P_d) for k in DAY_OUTPUT_FIELDS])
        self._c(self._lib.wg_get_day_outputs(self._h, member, C.byref(s)))
        return out

    def station_discharge(self, max_days=0):
        ns = max(self.n_stations, 0)
        total = np.zeros((self.n_members, ns))
        daily = np.zeros((self.n_members, max_days, ns)) if max_days else None
        self._c(self._lib.wg_get_station_discharge(self._h, _abi.ptr(total, _abi.P_d),
                                                   _abi.ptr(daily, _abi.P_d) if daily is not None else None,
                                                   max_days))
        return total, daily

    def station_daily_window(self):
        """(first day, days) of the kept daily station series, counted from the
        first day of the last step call."""
        a, b = C.c_int32(), C.c_int32()
        self._c(self._lib.wg_station_daily_window(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def reset_station_discharge(self):
        self._c(self._lib.wg_reset_station_discharge(self._h))

    def pot_cell_runoff_output(self, enable=True):
        """Keep G_potCellRunoff (wg_set_pot_cell_runoff): enabling when it is on
        restarts the sums at 0."""
        self._c(self._lib.wg_set_pot_cell_runoff(self._h, 1 if enable else 0))

    def pot_cell_runoff(self, member=0):
        """G_potCellRunoff of one member (km3 per cell, reference cell order),
        summed since the output was enabled or the last annual init."""
        out = np.zeros(self.n)
        self._c(self._lib.wg_get_pot_cell_runoff(self._h, int(member), _abi.ptr(out, _abi.P_d)))
        return out

    def set_water_use(self, nus_m3, nug_m3, withdrawal_irrig_m3, consumptive_irrig_m3, frac_return_gw,
                      year0_day=0, delayed=0, statics=None):
        """Water use (wg_set_water_use, a handle made with subtract_use 2 or 3):
        tables [n_years][N][12] in m3 per month, model year y starting at forcing
        day year0_day + 365 y; statics: the World of the last annual init (default:
        the model's own)."""
        tabs = [np.ascontiguousarray(t, dtype=np.float64) for t in
                (nus_m3, nug_m3, withdrawal_irrig_m3, consumptive_irrig_m3)]
        n_years = tabs[0].size // (12 * self.n)
        for t in tabs:
            if t.size != n_years * 12 * self.n or n_years == 0:
                raise ValueError("water-use tables must be [n_years][N][12]")
        fr = np.ascontiguousarray(frac_return_gw, dtype=np.float64)
        st = self._static if statics is None else (statics if isinstance(statics, _abi.StaticFields)
                                                  else World.static_fields(statics))
        self._use_static = st
        self._c(self._lib.wg_set_water_use(self._h, C.byref(st), int(n_years), int(year0_day),
                                           *[_abi.ptr(t, _abi.P_d) for t in tabs], _abi.ptr(fr, _abi.P_d),
                                           int(delayed)))

    def water_use_state(self, member=0):
        """[WG_USE_NFIELDS][N] water-use state of one member (fields _abi.USE_FIELDS)."""
        out = np.zeros((_abi.WG_USE_NFIELDS, self.n))
        self._c(self._lib.wg_get_water_use_state(self._h, int(member), _abi.ptr(out, _abi.P_d)))
        return out

    def set_water_use_state(self, state, member=0):
        a = np.ascontiguousarray(state, dtype=np.float64).reshape(_abi.WG_USE_NFIELDS, self.n)
        self._c(self._lib.wg_set_water_use_state(self._h, int(member), _abi.ptr(a, _abi.P_d)))

    def synchronize(self):
        self._c(self._lib.wg_synchronize(self._h))


def reservoir_year(year_to_use, reservoir_area_full, stor_cap_full, start_year, reg_lake_status=None):
    """Parts 1-2 of annualInit (wg_reservoir_year): the area and capacity of the
    reservoirs and regulated lakes in operation in the year to use."""
    lib = _abi.load_library()
    af = np.ascontiguousarray(reservoir_area_full, dtype=np.float64)
    cf = np.ascontiguousarray(stor_cap_full, dtype=np.float64)
    sy = np.ascontiguousarray(start_year, dtype=np.int32)
    rl = None if reg_lake_status is None else np.ascontiguousarray(reg_lake_status, dtype=np.int8)
    area, cap = np.zeros(len(af)), np.zeros(len(af))
    _check(lib.wg_reservoir_year(len(af), int(year_to_use), _abi.ptr(af, _abi.P_d), _abi.ptr(cf, _abi.P_d),
                                 _abi.ptr(sy, _abi.P_i32), None if rl is None else _abi.ptr(rl, _abi.P_i8),
                                 _abi.ptr(area, _abi.P_d), _abi.ptr(cap, _abi.P_d)))
    return area, cap
