"""The oracle with the deterministic exp/pow/log shared with the kernels against
the oracle built with the host C library's (glibc, the reference's own libm:
daily.cpp:251, 423, 859, 1126; routing.cpp:1940-1942, 2426, 2728, 3477-3478).

The GPU equals the shared-libm oracle bit for bit; these runs tie that oracle to
glibc arithmetic.  Every state field must agree to the parity tolerance of
tests/parity.py (1e-10 relative plus the fields' absolute floors).  Cells whose
land-area fraction reaches 0 are where a 1-ulp difference could be amplified
(routing.cpp:5343-5352 subtracts percent-scaled fractions there); the set is
computed and reported, and no cell, in it or outside it, is excluded.
"""
import numpy as np
import pytest

from parity import compare


def run_both(wg, oracle, world, days, seed, threads=8):
    f = world.forcing(seed=seed, day0=0, n_days=days)
    a = oracle.Oracle(world, "wg")
    b = oracle.Oracle(world, "glibc")
    a.set_threads(threads)
    b.set_threads(threads)
    min_laf = np.full(world.n, np.inf)
    for d in range(days):
        a.step_day(f[d, 0], f[d, 1], f[d, 2], f[d, 3], d + 1)
        b.step_day(f[d, 0], f[d, 1], f[d, 2], f[d, 3], d + 1)
        min_laf = np.minimum(min_laf, np.minimum(a.get_state()["laf"], b.get_state()["laf"]))
    return a, b, min_laf


@pytest.mark.parametrize("case", ["regional_365", "edge_365", "global_20"])
def test_shared_libm_oracle_matches_glibc_oracle(wg, oracle, case):
    if case == "regional_365":
        world, days, seed = wg.World(wg.regional_spec(seed=0)), 365, 0
    elif case == "edge_365":
        world, days, seed = wg.World(wg.regional_spec(seed=5, nrows=30, ncols=40, edge_cases=True,
                                                      glolake_every=20, res_every=25)), 365, 0
    else:
        world, days, seed = wg.World(wg.global_spec(seed=1)), 20, 1
    a, b, min_laf = run_both(wg, oracle, world, days, seed)
    zero_laf = int((min_laf <= 0).sum())
    print(f"{case}: {world.n} cells, {zero_laf} reach a land-area fraction of 0")
    bad = compare(a.get_state(), b.get_state())
    assert not bad, bad[:8]
    bad = compare(a.day_outputs(), b.day_outputs())
    assert not bad, bad[:8]
