"""The C ABI from a compiled C++ caller (tests/abi_driver.cpp).

The CPU test compiles the driver with g++ against include/watergap2_amd.h, links
it to the library and runs its host-only checks (routing schedule, .UNF bytes,
error codes).  The GPU test runs the prebuilt driver for 10 days of the 180-cell
regional grid and compares every state field and the station discharge with the
oracle at the parity tolerance of tests/parity.py.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from parity import assert_close

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "abi_driver.cpp")
PREBUILT = os.path.join(ROOT, "_build", "abi_driver")
LIBDIR = os.path.join(ROOT, "watergap2_amd")


def compile_driver(out):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ is not installed")
    subprocess.run([cxx, "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    "-o", out, SRC, "-L" + LIBDIR, "-lwatergap2_amd", "-Wl,-rpath," + LIBDIR],
                   check=True, capture_output=True, text=True)
    return out


def test_driver_compiles_links_and_checks(tmp_path, wg):
    wg._abi.load_library()  # the library is built
    exe = compile_driver(str(tmp_path / "abi_driver"))
    r = subprocess.run([exe, "check"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "abi_driver check ok (180 cells)" in r.stdout


@pytest.mark.gpu
def test_driver_ten_days_against_oracle(tmp_path, wg, oracle):
    exe = PREBUILT if os.path.exists(PREBUILT) else compile_driver(str(tmp_path / "abi_driver"))
    days = 10
    print(f"case: abi_driver run, regional 12 x 15 seed 0, {days} days", flush=True)
    r = subprocess.run([exe, "run", str(tmp_path), str(days)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    world = wg.World(wg.regional_spec(seed=0, nrows=12, ncols=15, glolake_every=10, res_every=20))
    n = world.n
    o = oracle.Oracle(world)
    o.set_stations(world.basin_outlet)
    o.run(world.forcing(seed=0, day0=0, n_days=days), 1)
    ref = o.get_state()
    got = {}
    for name, _, dtype, ch in wg._abi.STATE_FIELDS:
        a = np.fromfile(str(tmp_path / (name + ".bin")), dtype=dtype)
        assert a.size == n * ch, name
        got[name] = a.reshape(ref[name].shape)
    assert int(np.fromfile(str(tmp_path / "started.bin"), dtype=np.int32)[0]) == 1
    assert_close(got, ref, label="driver:")
    q = np.fromfile(str(tmp_path / "discharge.bin"), dtype=np.float64)
    assert_close({"discharge": q}, {"discharge": o.station_discharge()}, label="driver station:")
