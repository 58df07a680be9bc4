"""Restart files of the reference (text variants) for a model's state:
WGHM water storage states, snow in elevation and the additional out- and input
file (wghmStateFile.cpp, snowInElevationFile.cpp, additionalOutputInputFile.cpp),
with the conversions between a member's state and the files' values that the
reference's setStorages and its daily saving apply (daily.cpp:1896-1924,
routing.cpp:851-882, 4998-5020, integrateWGHM.cpp:830-846).  The work runs in the
library (wg_statefiles.cpp); this module marshals arrays.
"""
import os

import numpy as np

from . import _abi
from ._abi import NELEV, P_d

NSTORE = 10
NADDITIONAL = 53
STORE_NAMES = ["canopy", "snow", "soil", "loc_lake", "loc_wetl", "glo_lake", "glo_wetl", "reservoir", "river",
               "groundwater"]


def _lib():
    return _abi.load_library()


def _check(rc, what=""):
    if rc != 0:
        from . import WaterGapError
        raise WaterGapError(rc, what)


def _write(fn, path, arr, cols):
    arr = np.ascontiguousarray(arr, dtype=np.float64)
    n = arr.size // cols
    _check(getattr(_lib(), fn)(os.fsencode(path), n, _abi.ptr(arr, P_d)), path)


def _read(fn, path, n, cols):
    out = np.zeros((n, cols))
    _check(getattr(_lib(), fn)(os.fsencode(path), n, _abi.ptr(out, P_d)), path)
    return out


def write_wghm_state(path, storages_mm):
    _write("wg_write_wghm_state_txt", path, storages_mm, NSTORE)


def read_wghm_state(path, n):
    return _read("wg_read_wghm_state_txt", path, n, NSTORE)


def write_snow_elevation(path, snow_mm):
    _write("wg_write_snow_elevation_txt", path, snow_mm, NELEV)


def read_snow_elevation(path, n):
    return _read("wg_read_snow_elevation_txt", path, n, NELEV)


def write_additional(path, cols):
    _write("wg_write_additional_txt", path, cols, NADDITIONAL)


def read_additional(path, n):
    return _read("wg_read_additional_txt", path, n, NADDITIONAL)


def state_to_files(world, state):
    """(storages [N][10] mm, snow [N][101] mm, additional [N][53]) of a state."""
    from . import state_struct
    n = world.n
    st = world.static_fields()
    sto, snow, add = np.zeros((n, NSTORE)), np.zeros((n, NELEV)), np.zeros((n, NADDITIONAL))
    s = state_struct(state)
    _check(_lib().wg_state_to_files(n, st, s, _abi.ptr(sto, P_d), _abi.ptr(snow, P_d), _abi.ptr(add, P_d)))
    return sto, snow, add


def state_from_files(world, storages_mm, snow_mm, additional, base=None):
    """A state with the files' values; fields the files do not hold (the river's
    reduction factor and area fraction, the storage transfer, the started flag)
    come from `base` (a state dict) or stay 0."""
    from . import new_state, state_struct
    n = world.n
    state = {k: np.array(v, copy=True) if isinstance(v, np.ndarray) else v
             for k, v in (base.items() if base is not None else new_state(n).items())}
    s = state_struct(state)
    arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (storages_mm, snow_mm, additional)]
    _check(_lib().wg_state_from_files(n, world.static_fields(), *[_abi.ptr(a, P_d) for a in arrs], s))
    return state


def save_state(directory, world, state, tag="restart"):
    """Writes the three restart files of a state into a directory."""
    os.makedirs(directory, exist_ok=True)
    sto, snow, add = state_to_files(world, state)
    write_wghm_state(os.path.join(directory, f"wghm_state_{tag}.txt"), sto)
    write_snow_elevation(os.path.join(directory, f"snow_in_elevation_{tag}.txt"), snow)
    write_additional(os.path.join(directory, f"additional_{tag}.txt"), add)


def load_state(directory, world, tag="restart", base=None):
    n = world.n
    return state_from_files(world, read_wghm_state(os.path.join(directory, f"wghm_state_{tag}.txt"), n),
                            read_snow_elevation(os.path.join(directory, f"snow_in_elevation_{tag}.txt"), n),
                            read_additional(os.path.join(directory, f"additional_{tag}.txt"), n), base=base)
