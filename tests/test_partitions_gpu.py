"""The day loop in two partitions (the deepest routing groups on a stream of
their own) gives the bits of one partition: global grid, two members, daily
station series, potential cell runoff, other launch configurations, and the eager
path, which runs the whole grid's tables."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(wg, w, f, partitions, cfg=None, profiled=False):
    if partitions:
        os.environ.pop("WG_PARTITIONS", None)
    else:
        os.environ["WG_PARTITIONS"] = "1"
    try:
        M = 2
        rng = np.random.default_rng(1)
        gamma = np.tile(w.params[0], (M, 1)) * rng.uniform(0.8, 1.2, (M, w.n))
        cfa = rng.uniform(0.8, 1.2, (M, w.n))
        smax = np.tile(w.smax, (M, 1)).astype(np.float32)
        m = wg.WaterBalanceModel(w, n_members=M)
    finally:
        os.environ.pop("WG_PARTITIONS", None)
    if cfg:
        m.set_launch_config(*cfg)
    m.set_member_params(gamma, cfa, smax)
    m.set_stations(w.basin_outlet)
    m.pot_cell_runoff_output(True)
    m.load_forcing(f)
    m.step_days(0, 1, 25, profiled=profiled)
    m.step_days(25, 26, 15, profiled=profiled)
    tot, daily = m.station_discharge(max_days=15)
    out = {"parts": m.partitions(), "s0": m.get_state(0), "s1": m.get_state(1), "tot": tot, "daily": daily,
           "pot0": m.pot_cell_runoff(0), "pot1": m.pot_cell_runoff(1)}
    m.close()
    return out


def same(a, b, label):
    for k in ("s0", "s1"):
        for key in a[k]:
            assert np.array_equal(np.asarray(a[k][key]), np.asarray(b[k][key])), (label, k, key)
    for k in ("tot", "daily", "pot0", "pot1"):
        assert np.array_equal(a[k], b[k]), (label, k)


def test_partitions_change_no_result(wg, global_world):
    w = global_world
    f = w.forcing(seed=2, day0=0, n_days=40)
    print("case: global grid, one partition, 40 days, 2 members", flush=True)
    one = run(wg, w, f, False)
    assert len(one["parts"]) == 1 and one["parts"][0][0] == w.n
    print("case: global grid, two partitions", flush=True)
    two = run(wg, w, f, True)
    assert len(two["parts"]) == 2 and sum(c for c, _ in two["parts"]) == w.n
    assert two["parts"][0][1] == w.max_level + 1 and two["parts"][1][1] < two["parts"][0][1]
    same(one, two, "two partitions")
    for cfg in ((128, 256, 10, 3), (256, 256, 5, 2), (256, 128, 10, 0)):
        print(f"case: two partitions, launch {cfg}", flush=True)
        same(one, run(wg, w, f, True, cfg), f"launch {cfg}")
    print("case: two partitions, eager path", flush=True)
    same(one, run(wg, w, f, True, profiled=True), "eager")


def test_many_members_keep_one_partition(wg, global_world):
    """A day of many members is bandwidth-bound: no partitions above 8 members."""
    print("case: partitions by member count", flush=True)
    for M, parts in ((1, 2), (8, 2), (9, 1), (64, 1)):
        m = wg.WaterBalanceModel(global_world, n_members=M)
        assert len(m.partitions()) == parts, M
        m.close()
