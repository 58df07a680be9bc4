"""ctypes wrapper of the CPU oracle (oracle/libwg_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from watergap2_amd import _abi
from watergap2_amd._abi import NELEV, STATE_FIELDS, DAY_OUTPUT_FIELDS

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwg_oracle.so")
GLIBC_LIB_PATH = os.path.join(HERE, "libwg_oracle_glibc.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_LIBS = {}


def load(variant="wg"):
    """The oracle library: variant "wg" uses the deterministic exp/pow shared with
    the kernels, "glibc" the host C library's; any other value is the path of a
    library built from a modified copy of wg_oracle.c (the mutation tests)."""
    if variant in _LIBS:
        return _LIBS[variant]
    path = {"wg": LIB_PATH, "glibc": GLIBC_LIB_PATH}.get(variant, variant)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    i32 = C.c_int32
    P_i32, P_i16, P_i8, P_f, P_d = _abi.P_i32, _abi.P_i16, _abi.P_i8, _abi.P_f, _abi.P_d
    lib.wgo_inflow_outflow.argtypes = [i32, P_i8, P_i16, P_i16, P_i32, P_i32, P_i32]
    lib.wgo_rout_order.argtypes = [i32, P_i32, P_i32, P_i32, P_i32]
    lib.wgo_create.argtypes = [i32, C.POINTER(_abi.StaticFields)]
    lib.wgo_create.restype = C.c_void_p
    lib.wgo_destroy.argtypes = [C.c_void_p]
    lib.wgo_destroy.restype = None
    lib.wgo_annual_init.argtypes = [C.c_void_p, C.POINTER(_abi.StaticFields), C.POINTER(C.c_int32), C.c_int,
                                    C.POINTER(C.c_double)]
    lib.wgo_annual_init.restype = C.c_int
    lib.wgo_set_params.argtypes = [C.c_void_p, P_d, P_d, P_f]
    lib.wgo_set_params.restype = None
    lib.wgo_step_day.argtypes = [C.c_void_p, P_f, P_f, P_f, P_f, C.c_int, C.c_int]
    lib.wgo_step_day.restype = None
    lib.wgo_run.argtypes = [C.c_void_p, P_f, i32, i32]
    lib.wgo_run.restype = None
    lib.wgo_get_state.argtypes = [C.c_void_p, C.POINTER(_abi.StateStruct)]
    lib.wgo_set_state.argtypes = [C.c_void_p, C.POINTER(_abi.StateStruct)]
    lib.wgo_get_day_outputs.argtypes = [C.c_void_p, C.POINTER(_abi.DayOutputs)]
    lib.wgo_set_stations.argtypes = [C.c_void_p, i32, P_i32]
    lib.wgo_set_stations.restype = None
    lib.wgo_station_discharge.argtypes = [C.c_void_p, P_d]
    lib.wgo_station_discharge.restype = None
    lib.wgo_reset_station_discharge.argtypes = [C.c_void_p]
    lib.wgo_set_reference_scans.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.wgo_set_reference_scans.restype = None
    lib.wgo_reset_station_discharge.restype = None
    lib.wgo_pot_cell_runoff.argtypes = [C.c_void_p, P_d]
    lib.wgo_pot_cell_runoff.restype = None
    lib.wgo_reset_pot_cell_runoff.argtypes = [C.c_void_p]
    lib.wgo_reset_pot_cell_runoff.restype = None
    lib.wgo_set_water_use.argtypes = [C.c_void_p, C.c_int, P_d]
    lib.wgo_set_water_use.restype = C.c_int
    lib.wgo_water_use_year.argtypes = [C.c_void_p, P_d, P_d, P_d, P_d]
    lib.wgo_water_use_year.restype = None
    lib.wgo_water_use_year_end.argtypes = [C.c_void_p]
    lib.wgo_water_use_year_end.restype = None
    lib.wgo_get_water_use_state.argtypes = [C.c_void_p, P_d]
    lib.wgo_get_water_use_state.restype = None
    lib.wgo_set_water_use_state.argtypes = [C.c_void_p, P_d]
    lib.wgo_set_water_use_state.restype = None
    lib.wgo_perturb_forcing.argtypes = [C.c_uint64, i32, i32, C.c_double, C.c_double, P_f, P_f, P_f, P_f]
    lib.wgo_perturb_forcing.restype = None
    lib.wgo_month_of_day.argtypes = [C.c_int]
    lib.wgo_set_threads.argtypes = [C.c_void_p, C.c_int]
    lib.wgo_set_threads.restype = None
    lib.wgo_normal_pair.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, P_d]
    lib.wgo_normal_pair.restype = None
    lib.wgo_sincos2pi.argtypes = [C.c_double, P_d, P_d]
    lib.wgo_sincos2pi.restype = None
    lib.wgo_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.wgo_philox.restype = None
    for fn in ("wgo_exp", "wgo_log"):
        getattr(lib, fn).argtypes = [C.c_double]
        getattr(lib, fn).restype = C.c_double
    lib.wgo_pow.argtypes = [C.c_double, C.c_double]
    lib.wgo_pow.restype = C.c_double
    P_u16 = C.POINTER(C.c_uint16)
    lib.wgo_flow_accumulation.argtypes = [i32, P_i32, P_i16]
    lib.wgo_derive_basins.argtypes = [i32, P_i8, P_i32, P_u16, P_u16, P_i32]
    lib.wgo_reindex_basins.argtypes = [i32, i32, P_u16, P_i32]
    lib.wgo_cell_distances.argtypes = [P_f]
    lib.wgo_cell_distances.restype = None
    lib.wgo_river_slope.argtypes = [i32, P_f, P_f, P_f, P_i32, P_i8, P_i16, P_f]
    lib.wgo_river_slope.restype = None
    lib.wgo_river_length.argtypes = [i32, P_f, P_i16, P_i16, P_f, P_i32, P_f]
    lib.wgo_river_length.restype = None
    lib.wgo_alloc_coeff.argtypes = [i32, P_i32, P_d, P_d, P_d]
    lib.wgo_alloc_coeff.restype = None
    lib.wgo_reservoir_start_month.argtypes = [i32, P_d, P_d, P_i8]
    lib.wgo_reservoir_start_month.restype = None
    _LIBS[variant] = lib
    return lib


def build_variant(replacements, out_path, glibc=False):
    """Builds the oracle from a copy of wg_oracle.c with each (old, new) text
    replacement applied (every `old` must occur exactly once) into out_path: the
    mutants that show a pinning check fails when the arithmetic it pins is wrong."""
    src = open(os.path.join(HERE, "wg_oracle.c")).read()
    for old, new in replacements:
        if src.count(old) != 1:
            raise ValueError(f"replacement target occurs {src.count(old)} times: {old!r}")
        src = src.replace(old, new)
    c_path = out_path[:-3] + ".c"
    with open(c_path, "w") as f:
        f.write(src)
    cmd = ["gcc", "-O2", "-std=c99", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-w",
           "-I", HERE, "-shared", "-o", out_path, c_path, os.path.join(HERE, "wg_prepare_oracle.c"), "-lm"]
    if glibc:
        cmd.insert(1, "-DWG_ORACLE_GLIBC")
    subprocess.run(cmd, check=True)
    return out_path


def _p(arr, ct):
    return _abi.ptr(arr, ct)


def flow_accumulation(inflow9):
    """rout_prepare.cpp:441-494 with the reference's sweeps."""
    inflow9 = np.ascontiguousarray(inflow9, dtype=np.int32)
    acc = np.zeros(len(inflow9), np.int16)
    load().wgo_flow_accumulation(len(inflow9), _p(inflow9, _abi.P_i32), _p(acc, _abi.P_i16))
    return acc


def derive_basins(ldd, inflow9):
    """rout_prepare.cpp:497-604: (ldd after the no-data fix, basins, cells to the
    outlet, number of endpoints, basins renumbered, number of multi-cell basins)."""
    P_u16 = C.POINTER(C.c_uint16)
    ldd = np.array(ldd, dtype=np.int8)
    inflow9 = np.ascontiguousarray(inflow9, dtype=np.int32)
    n = len(ldd)
    ws = np.zeros(n, np.uint16)
    dist = np.zeros(n, np.uint16)
    nb = C.c_int32()
    lib = load()
    lib.wgo_derive_basins(n, _p(ldd, _abi.P_i8), _p(inflow9, _abi.P_i32), _p(ws, P_u16), _p(dist, P_u16), C.byref(nb))
    ws2 = ws.copy()
    nm = C.c_int32()
    lib.wgo_reindex_basins(n, nb.value, _p(ws2, P_u16), C.byref(nm))
    return ldd, ws, dist, nb.value, ws2, nm.value


def cell_distances():
    d = np.zeros(9 * 360, np.float32)
    load().wgo_cell_distances(_p(d, _abi.P_f))
    return d


def river_slope(meander, dist, altitude, outflow, ldd, row):
    args = [np.ascontiguousarray(a, dtype=t) for a, t in ((meander, np.float32), (dist, np.float32),
                                                           (altitude, np.float32), (outflow, np.int32),
                                                           (ldd, np.int8), (row, np.int16))]
    out = np.zeros(len(args[0]), np.float32)
    cts = [_abi.P_f, _abi.P_f, _abi.P_f, _abi.P_i32, _abi.P_i8, _abi.P_i16]
    load().wgo_river_slope(len(out), *[_p(a, c) for a, c in zip(args, cts)], _p(out, _abi.P_f))
    return out


def river_length(meander, row, col, dist, outflow):
    args = [np.ascontiguousarray(a, dtype=t) for a, t in ((meander, np.float32), (row, np.int16), (col, np.int16),
                                                           (dist, np.float32), (outflow, np.int32))]
    out = np.zeros(len(args[0]), np.float32)
    cts = [_abi.P_f, _abi.P_i16, _abi.P_i16, _abi.P_f, _abi.P_i32]
    load().wgo_river_length(len(out), *[_p(a, c) for a, c in zip(args, cts)], _p(out, _abi.P_f))
    return out


def alloc_coeff(outflow, reservoir_area, mean_outflow):
    outflow = np.ascontiguousarray(outflow, dtype=np.int32)
    ra = np.ascontiguousarray(reservoir_area, dtype=np.float64)
    mo = np.ascontiguousarray(mean_outflow, dtype=np.float64)
    out = np.zeros((len(outflow), 5))
    load().wgo_alloc_coeff(len(outflow), _p(outflow, _abi.P_i32), _p(ra, _abi.P_d), _p(mo, _abi.P_d),
                           _p(out, _abi.P_d))
    return out


def reservoir_start_month(monthly, annual):
    monthly = np.ascontiguousarray(monthly, dtype=np.float64)
    annual = np.ascontiguousarray(annual, dtype=np.float64)
    out = np.zeros(len(annual), np.int8)
    load().wgo_reservoir_start_month(len(annual), _p(monthly, _abi.P_d), _p(annual, _abi.P_d), _p(out, _abi.P_i8))
    return out


def rout_order(inflow9, outflow):
    lib = load()
    inflow9 = np.ascontiguousarray(inflow9, dtype=np.int32)
    outflow = np.ascontiguousarray(outflow, dtype=np.int32)
    n = len(outflow)
    order = np.zeros(n, dtype=np.int32)
    level = np.zeros(n, dtype=np.int32)
    rc = lib.wgo_rout_order(n, _abi.ptr(inflow9, _abi.P_i32), _abi.ptr(outflow, _abi.P_i32),
                            _abi.ptr(order, _abi.P_i32), _abi.ptr(level, _abi.P_i32))
    return rc, order, level


def inflow_outflow(world):
    lib = load()
    n = world.n
    inflow9 = np.zeros((n, 9), dtype=np.int32)
    outflow = np.zeros(n, dtype=np.int32)
    lib.wgo_inflow_outflow(n, _abi.ptr(world.ldd, _abi.P_i8), _abi.ptr(world.row, _abi.P_i16),
                           _abi.ptr(world.col, _abi.P_i16), _abi.ptr(world.cell_at, _abi.P_i32),
                           _abi.ptr(inflow9, _abi.P_i32), _abi.ptr(outflow, _abi.P_i32))
    return inflow9, outflow


def philox(ctr, key):
    """One Philox4x32-10 block of the oracle's generator."""
    c = (C.c_uint32 * 4)(*[int(v) for v in ctr])
    k = (C.c_uint32 * 2)(*[int(v) for v in key])
    load().wgo_philox(c, k)
    return [int(v) for v in c]


def normal_pairs(seed, cells, day):
    """[len(cells), 2] normal deviates of the perturbation for one member and day."""
    lib = load()
    out = np.zeros((len(cells), 2))
    buf = (C.c_double * 2)()
    for k, c in enumerate(cells):
        lib.wgo_normal_pair(int(seed), int(c), int(day), C.cast(buf, _abi.P_d))
        out[k] = buf[0], buf[1]
    return out


def sincos2pi(u):
    lib = load()
    s, c = C.c_double(), C.c_double()
    lib.wgo_sincos2pi(float(u), C.byref(s), C.byref(c))
    return s.value, c.value


def perturb_forcing(seed, day_index, temp, prec, prec_sigma, temp_sigma):
    lib = load()
    temp = np.ascontiguousarray(temp, dtype=np.float32)
    prec = np.ascontiguousarray(prec, dtype=np.float32)
    t_out = np.empty_like(temp)
    p_out = np.empty_like(prec)
    lib.wgo_perturb_forcing(int(seed), int(day_index), len(temp), prec_sigma, temp_sigma,
                            _abi.ptr(temp, _abi.P_f), _abi.ptr(prec, _abi.P_f),
                            _abi.ptr(t_out, _abi.P_f), _abi.ptr(p_out, _abi.P_f))
    return t_out, p_out


class Oracle:
    """One CPU model: calcNewDay for every cell, serial routing, land-area update."""

    def __init__(self, world, variant="wg"):
        self._lib = load(variant)
        self.world = world
        self.n = world.n
        self._static = world.static_fields()
        self._m = self._lib.wgo_create(self.n, C.byref(self._static))
        self.n_stations = 0

    def __del__(self):
        try:
            if self._m:
                self._lib.wgo_destroy(self._m)
                self._m = None
        except Exception:
            pass

    def annual_init(self, year_statics, outflow_cell=None, apply_changes=True):
        """wgo_annual_init: routingClass::annualInit for a later year."""
        import watergap2_amd as _wg
        if not hasattr(self, "_glo_res_prev"):
            self._glo_res_prev = np.ascontiguousarray(self.world.glo_res, dtype=np.float64).copy()
        self._year_st = _wg.World.static_fields(year_statics)   # kept: the oracle reads it by pointer
        self._year_obj = year_statics
        oc = None if outflow_cell is None else np.ascontiguousarray(outflow_cell, dtype=np.int32)
        self._year_oc = oc
        rc = self._lib.wgo_annual_init(self._m, C.byref(self._year_st), None if oc is None else _p(oc, C.POINTER(C.c_int32)),
                                       1 if apply_changes else 0, _p(self._glo_res_prev, C.POINTER(C.c_double)))
        if rc != 0:
            raise RuntimeError(f"wgo_annual_init returned {rc}")

    def set_reference_scans(self, station_scan=False, aggr_scan=False):
        self._lib.wgo_set_reference_scans(self._m, int(station_scan), int(aggr_scan))

    def set_threads(self, threads):
        self._lib.wgo_set_threads(self._m, int(threads))

    def set_params(self, gamma=None, cfa=None, smax=None):
        g = None if gamma is None else np.ascontiguousarray(gamma, dtype=np.float64)
        c = None if cfa is None else np.ascontiguousarray(cfa, dtype=np.float64)
        s = None if smax is None else np.ascontiguousarray(smax, dtype=np.float32)
        self._keep = (g, c, s)
        self._lib.wgo_set_params(self._m, _abi.ptr(g, _abi.P_d), _abi.ptr(c, _abi.P_d),
                                 _abi.ptr(s, _abi.P_f))

    def run(self, forcing, first_doy=1):
        forcing = np.ascontiguousarray(forcing, dtype=np.float32)
        self._lib.wgo_run(self._m, _abi.ptr(forcing, _abi.P_f), forcing.shape[0], first_doy)

    def step_day(self, temp, prec, sw, lw, doy):
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (temp, prec, sw, lw)]
        month = self._lib.wgo_month_of_day(doy)
        self._lib.wgo_step_day(self._m, *[_abi.ptr(a, _abi.P_f) for a in arrs], doy, month)

    def get_state(self):
        from watergap2_amd import new_state, state_struct
        st = new_state(self.n)
        s = state_struct(st)
        self._lib.wgo_get_state(self._m, C.byref(s))
        st["started"] = int(s.started)
        return st

    def set_state(self, st):
        from watergap2_amd import new_state, state_struct
        full = new_state(self.n)
        for k in full:
            if k in st and k != "started":
                full[k] = np.ascontiguousarray(st[k], dtype=full[k].dtype).reshape(full[k].shape)
        full["started"] = st.get("started", 0)
        self._lib.wgo_set_state(self._m, C.byref(state_struct(full)))

    def day_outputs(self):
        out = {k: np.zeros(self.n) for k in DAY_OUTPUT_FIELDS}
        s = _abi.DayOutputs(*[_abi.ptr(out[k], _abi.P_d) for k in DAY_OUTPUT_FIELDS])
        self._lib.wgo_get_day_outputs(self._m, C.byref(s))
        return out

    def set_stations(self, cells):
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.n_stations = len(cells)
        self._lib.wgo_set_stations(self._m, len(cells), _abi.ptr(cells, _abi.P_i32))

    def station_discharge(self):
        out = np.zeros(self.n_stations)
        self._lib.wgo_station_discharge(self._m, _abi.ptr(out, _abi.P_d))
        return out

    def reset_station_discharge(self):
        self._lib.wgo_reset_station_discharge(self._m)

    def pot_cell_runoff(self):
        out = np.zeros(self.n)
        self._lib.wgo_pot_cell_runoff(self._m, _abi.ptr(out, _abi.P_d))
        return out

    def reset_pot_cell_runoff(self):
        self._lib.wgo_reset_pot_cell_runoff(self._m)

    # water use (subtract_use 2/3, use_alloc 0, aggrNUsGloLakResOpt 0)
    def set_water_use(self, frac_return_gw, delayed=0):
        self._use_frgi = np.ascontiguousarray(frac_return_gw, dtype=np.float64)
        rc = self._lib.wgo_set_water_use(self._m, int(delayed), _abi.ptr(self._use_frgi, _abi.P_d))
        if rc:
            raise MemoryError("wgo_set_water_use")

    def water_use_year(self, nus_m3, nug_m3, wu_m3, cu_m3):
        """The year's [N][12] tables in m3 per month (dailyNUInit)."""
        a = [np.ascontiguousarray(x, dtype=np.float64).reshape(self.n, 12) for x in (nus_m3, nug_m3, wu_m3, cu_m3)]
        self._lib.wgo_water_use_year(self._m, *[_abi.ptr(x, _abi.P_d) for x in a])

    def water_use_year_end(self):
        self._lib.wgo_water_use_year_end(self._m)

    def water_use_state(self):
        out = np.zeros((_abi.WG_USE_NFIELDS, self.n))
        self._lib.wgo_get_water_use_state(self._m, _abi.ptr(out, _abi.P_d))
        return out

    def set_water_use_state(self, state):
        a = np.ascontiguousarray(state, dtype=np.float64).reshape(_abi.WG_USE_NFIELDS, self.n)
        self._lib.wgo_set_water_use_state(self._m, _abi.ptr(a, _abi.P_d))
