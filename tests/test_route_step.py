"""The chain kernels route most cells with a split step: step_pre (the arid
groundwater step, which needs no upstream water) once per batch, then step_fast
(lake, wetland and river with selects in place of branches); cells that step_fast
flags take the general chain_step (routing.cpp:2623-3958).  Both are in
watergap2_amd/csrc/wg_route_step.h, which this test compiles for the CPU with the
kernels' floating-point flags and runs on random routing records covering every
flag combination, dry lakes, wetlands and rivers, and every day of the year: the
split step must give chain_step's bits, for every output it writes.
"""
import os
import shutil
import subprocess

import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "watergap2_amd", "csrc")

SRC = r"""
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
using std::fabs;
using std::fma;
#include "wg_device.h"
#include "wg_route_step.h"

static bool same(double a, double b) { return std::memcmp(&a, &b, 8) == 0; }

int main(int argc, char **argv) {
  if (argc > 2) {  // the calendar facts of every day of the year
    for (int doy = 1; doy <= 365; doy++) {
      const DayInfo d = day_info(doy);
      std::printf("%d %d %d %.17g\n", doy, d.month, int(d.first_of_month), d.days_in_month);
    }
    return 0;
  }
  const long n = std::atol(argv[1]);
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(0., 1.);
  long n_fast = 0, n_bad = 0, n_res = 0, n_dry = 0;
  for (long it = 0; it < n; it++) {
    ChainRec c;
    int flags = CFL_CONT;
    if (U(rng) < 0.3) flags |= CFL_LAKE;
    if (U(rng) < 0.1) flags |= CFL_RES;
    if (U(rng) < 0.6) flags |= CFL_GWET;
    if (U(rng) < 0.4) flags |= CFL_ARID;
    if (U(rng) < 0.9) flags |= CFL_LDD;
    flags |= (1 + int(U(rng) * 2)) << 8 | (1 + int(U(rng) * 12)) << 16;
    c.flags = flags;
    // magnitudes over three decades, with exact zeros
    auto r = [&](double s) { const double x = U(rng); return x < 0.05 ? 0. : s * std::pow(10., -3. * U(rng)) * x; };
    for (int f = 0; f < CF_COUNT; f++) c.v[f] = r(1.0);
    if (!(flags & CFL_GWET)) c.v[CF_GW_WP] = c.v[CF_GW_PET] = c.v[CF_GW_MAX] = c.v[CF_GW_PREV] = c.v[CF_GW_S0E] = c.v[CF_GW_GWR] = 0.;
    if (!(flags & CFL_ARID)) c.v[CF_A_PREV] = c.v[CF_A_S0E] = 0.;
    if (!(flags & CFL_LAKE)) c.v[CF_GL_GWR] = 0.;
    if (U(rng) < 0.2) c.v[CF_GW_PET] = 5. * U(rng);   // wetland running dry
    if (U(rng) < 0.2) c.v[CF_RV_EVAP] = 5. * U(rng);  // river running dry
    if (U(rng) < 0.2) c.v[CF_GL_USE] = 5. * U(rng);   // lake running dry
    if (U(rng) < 0.05) c.v[CF_RV_PREV] = 0.;
    c.inv_ksw = r(2.); c.ome_sw = U(rng); c.inv_kgw = r(2.); c.ome_gw = U(rng);
    c.area = 2500. * U(rng); c.cf100 = U(rng);
    const double up = U(rng) < 0.1 ? 0. : r(3.);
    const DayInfo d = day_info(1 + int(U(rng) * 365));
    const ChainOut a = chain_step(c, up, d);
    bool slow;
    const StepPre pre = step_pre(c);
    ChainOut b = step_fast(c, pre, up, slow);
    if (slow) b = chain_step(c, up, d);
    else n_fast++;
    if (flags & CFL_RES) n_res++;
    if (a.river == 0.) n_dry++;
    bool ok = same(a.tv, b.tv) && same(a.river, b.river) && same(a.gw, b.gw) && same(a.glo_wetl, b.glo_wetl);
    if (flags & CFL_LAKE) ok = ok && same(a.glo_lake, b.glo_lake);
    if (flags & CFL_RES) ok = ok && same(a.res, b.res) && same(a.k_release, b.k_release);
    if (!ok && n_bad++ < 5)
      std::printf("mismatch %ld flags %x: tv %.17g %.17g river %.17g %.17g\n", it, flags, a.tv, b.tv, a.river, b.river);
  }
  std::printf("%ld %ld %ld %ld %ld\n", n, n_fast, n_bad, n_res, n_dry);
  return 0;
}
"""


@pytest.fixture(scope="module")
def route_step_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    d = tmp_path_factory.mktemp("route_step")
    src, exe = str(d / "route_step.cpp"), str(d / "route_step")
    with open(src, "w") as f:
        f.write(SRC)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-I", CSRC, src, "-o", exe],
                   check=True)
    return exe


def test_day_info_matches_the_calendar(route_step_exe):
    """day_info's month, first-of-month flag and month length (the reservoir's
    release year starts on the first day of its start month, routing.cpp:2956-2957)
    for every day of a 365-day year, against Python's calendar."""
    import calendar
    import datetime
    out = subprocess.run([route_step_exe, "0", "days"], check=True, capture_output=True, text=True).stdout
    rows = [line.split() for line in out.strip().splitlines()]
    assert len(rows) == 365
    for doy, month, first, length in rows:
        date = datetime.date(2001, 1, 1) + datetime.timedelta(days=int(doy) - 1)
        assert int(month) == date.month - 1
        assert int(first) == int(date.day == 1)
        assert float(length) == calendar.monthrange(2001, date.month)[1]


def test_split_step_gives_chain_step_bits(route_step_exe):
    out = subprocess.run([route_step_exe, "1000000"], check=True, capture_output=True, text=True).stdout
    lines = out.strip().splitlines()
    n, n_fast, n_bad, n_res, n_dry = map(int, lines[-1].split())
    assert n_bad == 0, out
    # the draw covers both paths, reservoirs and dry rivers
    assert n_fast > n // 2 and n - n_fast > n // 20 and n_res > n // 20 and n_dry > n // 100, out
