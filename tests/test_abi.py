"""The C-ABI library loads and exports every entry point include/watergap2_amd.h declares."""
import os
import re

from watergap2_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "watergap2_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wg_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_the_abi():
    funcs = header_functions()
    assert sorted(_abi.ABI_SYMBOLS) == funcs


def test_library_exports_every_symbol():
    lib = _abi.load_library()
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_default_options_is_parity_profile():
    import ctypes as C
    lib = _abi.load_library()
    opt = _abi.OptionsProfile()
    lib.wg_default_options(C.byref(opt))
    assert (opt.time_series, opt.cloud, opt.intercept, opt.pet_opt, opt.river_velo_opt) == (1, 1, 1, 0, 2)
    assert (opt.subtract_use, opt.arid_area_opt, opt.fractional_routing_opt, opt.river_evapo_opt) == (0, 1, 1, 1)


def test_create_rejects_bad_arguments(wg, regional_world):
    import ctypes as C
    lib = _abi.load_library()
    h = C.c_void_p()
    s = regional_world.static_fields()
    opt = _abi.OptionsProfile()
    lib.wg_default_options(C.byref(opt))
    assert lib.wg_create(0, C.byref(s), C.byref(opt), 1, C.byref(h)) == -1
    assert lib.wg_create(regional_world.n, None, C.byref(opt), 1, C.byref(h)) == -1
    opt.pet_opt = 3  # Hargreaves is not part of the implemented profile
    assert lib.wg_create(regional_world.n, C.byref(s), C.byref(opt), 1, C.byref(h)) == -1
    lib.wg_default_options(C.byref(opt))
    bad = regional_world.aindex.copy()
    bad[7] = 2  # invalid arid index, daily.cpp:345-347
    s.aindex = _abi.ptr(bad, _abi.P_i16)
    assert lib.wg_create(regional_world.n, C.byref(s), C.byref(opt), 1, C.byref(h)) == -5
