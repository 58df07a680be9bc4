"""Routing preprocessing: the host side of prepare_routing_files
(rout_prepare.cpp:61-336) over the C ABI (watergap2_amd/csrc/wg_prepare.cpp).

:func:`prepare_routing` turns a flow-direction map on the 0.5-degree lattice into
the routing files the day loop reads (inflow and outflow cells, flow
accumulation, basins, river slope and length, routing order, and for reservoirs
the allocation coefficients and start months).  The single steps are exported as
well, under the reference's function names.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import P_d, P_f, P_i8, P_i16, P_i32

_P_u16 = C.POINTER(C.c_uint16)


def _lib():
    return _abi.load_library()


def _check(rc):
    if rc != 0:
        from . import WaterGapError
        raise WaterGapError(rc)


def _arr(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def ldd_from_flowdir(flowdir):
    """Arc/Info flow codes -> WaterGAP LDD (rout_prepare.cpp:68-112)."""
    flowdir = _arr(flowdir, np.int16)
    ldd = np.zeros(len(flowdir), np.int8)
    _check(_lib().wg_ldd_from_flowdir(len(flowdir), _abi.ptr(flowdir, P_i16), _abi.ptr(ldd, P_i8)))
    return ldd


def replace_flow_to_ocean(row, col, cell_at, ldd):
    """Cells flowing into the ocean become sinks (rout_prepare.cpp:287-336)."""
    row, col, cell_at = _arr(row, np.int16), _arr(col, np.int16), _arr(cell_at, np.int32)
    ldd = np.array(ldd, dtype=np.int8)
    _check(_lib().wg_replace_flow_to_ocean(len(ldd), _abi.ptr(row, P_i16), _abi.ptr(col, P_i16),
                                           _abi.ptr(cell_at, P_i32), _abi.ptr(ldd, P_i8)))
    return ldd


def flow_accumulation(inflow9):
    """Cells draining through each cell, itself included; 0 on loops (rout_prepare.cpp:441-494)."""
    inflow9 = _arr(inflow9, np.int32)
    acc = np.zeros(len(inflow9), np.int16)
    _check(_lib().wg_flow_accumulation(len(inflow9), _abi.ptr(inflow9, P_i32), _abi.ptr(acc, P_i16)))
    return acc


def derive_basins(ldd, inflow9):
    """Basins and cells to the outlet (rout_prepare.cpp:497-571), then the
    renumbering of the basins of several cells (573-604).  Returns (corrected LDD,
    basins, cells to outlet, endpoints, renumbered basins, multi-cell basins)."""
    ldd = np.array(ldd, dtype=np.int8)
    inflow9 = _arr(inflow9, np.int32)
    n = len(ldd)
    basin = np.zeros(n, np.uint16)
    dist = np.zeros(n, np.uint16)
    nb = C.c_int32()
    _check(_lib().wg_derive_basins(n, _abi.ptr(ldd, P_i8), _abi.ptr(inflow9, P_i32), _abi.ptr(basin, _P_u16),
                                   _abi.ptr(dist, _P_u16), C.byref(nb)))
    basin2 = basin.copy()
    nm = C.c_int32()
    _check(_lib().wg_reindex_basins(n, nb.value, _abi.ptr(basin2, _P_u16), C.byref(nm)))
    return ldd, basin, dist, nb.value, basin2, nm.value


def neighbour_cells(row, col, cell_at):
    row, col, cell_at = _arr(row, np.int16), _arr(col, np.int16), _arr(cell_at, np.int32)
    out = np.zeros((len(row), 8), np.int32)
    _check(_lib().wg_neighbour_cells(len(row), _abi.ptr(row, P_i16), _abi.ptr(col, P_i16), _abi.ptr(cell_at, P_i32),
                                     _abi.ptr(out, P_i32)))
    return out


def cell_distances():
    """[9 directions][360 rows] distances in km (rout_prepare.cpp:694-752)."""
    d = np.zeros(9 * 360, np.float32)
    _check(_lib().wg_cell_distances(_abi.ptr(d, P_f)))
    return d


def river_slope(meander, dist, altitude, outflow, ldd, row):
    a = [_arr(meander, np.float32), _arr(dist, np.float32), _arr(altitude, np.float32), _arr(outflow, np.int32),
         _arr(ldd, np.int8), _arr(row, np.int16)]
    out = np.zeros(len(a[0]), np.float32)
    cts = [P_f, P_f, P_f, P_i32, P_i8, P_i16]
    _check(_lib().wg_river_slope(len(out), *[_abi.ptr(x, c) for x, c in zip(a, cts)], _abi.ptr(out, P_f)))
    return out


def river_length(meander, row, col, dist, outflow):
    a = [_arr(meander, np.float32), _arr(row, np.int16), _arr(col, np.int16), _arr(dist, np.float32),
         _arr(outflow, np.int32)]
    out = np.zeros(len(a[0]), np.float32)
    cts = [P_f, P_i16, P_i16, P_f, P_i32]
    _check(_lib().wg_river_length(len(out), *[_abi.ptr(x, c) for x, c in zip(a, cts)], _abi.ptr(out, P_f)))
    return out


def alloc_coeff(outflow, reservoir_area, mean_outflow):
    outflow, ra, mo = _arr(outflow, np.int32), _arr(reservoir_area, np.float64), _arr(mean_outflow, np.float64)
    out = np.zeros((len(outflow), 5))
    _check(_lib().wg_alloc_coeff(len(outflow), _abi.ptr(outflow, P_i32), _abi.ptr(ra, P_d), _abi.ptr(mo, P_d),
                                 _abi.ptr(out, P_d)))
    return out


def reservoir_start_month(monthly, annual):
    monthly, annual = _arr(monthly, np.float64), _arr(annual, np.float64)
    out = np.zeros(len(annual), np.int8)
    _check(_lib().wg_reservoir_start_month(len(annual), _abi.ptr(monthly, P_d), _abi.ptr(annual, P_d),
                                           _abi.ptr(out, P_i8)))
    return out


def prepare_routing(row, col, cell_at, altitude, meander, flowdir=None, ldd=None, reservoir_area=None,
                    mean_outflow=None, mean_outflow_monthly=None):
    """prepare_routing_files (rout_prepare.cpp:61-336): returns a dict of the
    routing files named after the reference's outputs."""
    n = len(row)
    keep = dict(flowdir=_arr(flowdir, np.int16), ldd=_arr(ldd, np.int8), row=_arr(row, np.int16),
                col=_arr(col, np.int16), cell_at=_arr(cell_at, np.int32), altitude=_arr(altitude, np.float32),
                meander=_arr(meander, np.float32), reservoir_area=_arr(reservoir_area, np.float64),
                mean_outflow=_arr(mean_outflow, np.float64),
                mean_outflow_monthly=_arr(mean_outflow_monthly, np.float64))
    cts = dict(flowdir=P_i16, ldd=P_i8, row=P_i16, col=P_i16, cell_at=P_i32, altitude=P_f, meander=P_f,
               reservoir_area=P_d, mean_outflow=P_d, mean_outflow_monthly=P_d)
    ins = _abi.RoutingInputs(n_cells=n, **{k: _abi.ptr(v, cts[k]) for k, v in keep.items()})
    res = reservoir_area is not None
    out = dict(ldd=np.zeros(n, np.int8), inflow9=np.zeros((n, 9), np.int32), outflow=np.zeros(n, np.int32),
               flow_acc=np.zeros(n, np.int16), basins=np.zeros(n, np.uint16), cells_to_outlet=np.zeros(n, np.uint16),
               cell_dist=np.zeros(9 * 360, np.float32), river_slope=np.zeros(n, np.float32),
               river_length=np.zeros(n, np.float32), rout_order=np.zeros(n, np.int32),
               alloc_coeff=np.zeros((n, 5)) if res else None, start_month=np.zeros(n, np.int8) if res else None)
    octs = dict(ldd=P_i8, inflow9=P_i32, outflow=P_i32, flow_acc=P_i16, basins=_P_u16, cells_to_outlet=_P_u16,
                cell_dist=P_f, river_slope=P_f, river_length=P_f, rout_order=P_i32, alloc_coeff=P_d, start_month=P_i8)
    files = _abi.RoutingFiles(**{k: _abi.ptr(v, octs[k]) for k, v in out.items()}, n_basins=0)
    _check(_lib().wg_prepare_routing(C.byref(ins), C.byref(files)))
    out["n_basins"] = int(files.n_basins)
    return out
