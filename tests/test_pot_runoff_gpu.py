"""The potential cell runoff on the GPU (wg_set_pot_cell_runoff) against the
oracle: members with their own CFA and gamma, every chain kernel, a grid with
inland sinks, (semi-)arid cells, lakes and reservoirs; the output changes no
state; wg_annual_init restarts the sums."""
import numpy as np
import pytest

from parity import assert_close

pytestmark = pytest.mark.gpu


def sink_world(wg):
    return wg.World(wg.global_spec(seed=2, n_land=3000, n_outlets=30, n_sinks=15))


@pytest.mark.parametrize("mode", [3, 2, 1, 0])
def test_pot_cell_runoff_matches_oracle(wg, oracle, regional_edge_world, mode):
    for w, label in ((regional_edge_world, "edge"), (sink_world(wg), "sinks")):
        print(f"case: potential cell runoff, {label} world, chain mode {mode}, 2 members, 50 days", flush=True)
        M, n_days = 2, 50
        rng = np.random.default_rng(7)
        gamma = np.tile(w.params[0], (M, 1)) * rng.uniform(0.8, 1.2, (M, w.n))
        cfa = rng.uniform(0.7, 1.3, (M, w.n))
        smax = np.tile(w.smax, (M, 1)).astype(np.float32)
        f = w.forcing(seed=3, day0=0, n_days=n_days)
        m = wg.WaterBalanceModel(w, n_members=M)
        m.set_launch_config(256, 512 if mode == 3 else 256, 10, mode)
        m.set_member_params(gamma, cfa, smax)
        m.pot_cell_runoff_output(True)
        m.load_forcing(f)
        m.step_days(0, 1, 30)
        m.step_days(30, 31, n_days - 30)  # two calls: the sums run on
        for k in range(M):
            o = oracle.Oracle(w)
            o.set_params(gamma[k], cfa[k], smax[k])
            o.run(f, 1)
            pot = m.pot_cell_runoff(k)
            assert np.array_equal(pot, o.pot_cell_runoff()), (label, mode, k)
            assert_close(m.get_state(k), o.get_state(), label=f"{label} m{k}:")
        assert np.abs(m.pot_cell_runoff(0)).max() > 0


def test_output_changes_no_state_and_restarts(wg, oracle, regional_edge_world):
    print("case: potential cell runoff on/off, restart, 40 days", flush=True)
    w = regional_edge_world
    f = w.forcing(seed=4, day0=0, n_days=40)
    runs = []
    for on in (False, True):
        m = wg.WaterBalanceModel(w)
        if on:
            m.pot_cell_runoff_output(True)
        m.load_forcing(f)
        m.step_days(0, 1, 40)
        runs.append(m)
    a, b = runs[0].get_state(), runs[1].get_state()
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    with pytest.raises(RuntimeError):
        runs[0].pot_cell_runoff()
    m = runs[1]
    # enabling again restarts the sums; the next 10 days equal an oracle run's sums over them
    o = oracle.Oracle(w)
    o.run(f[:30], 1)
    m2 = wg.WaterBalanceModel(w)
    m2.pot_cell_runoff_output(True)
    m2.load_forcing(f)
    m2.step_days(0, 1, 30)
    m2.pot_cell_runoff_output(True)
    o.reset_pot_cell_runoff()
    m2.step_days(30, 31, 10)
    o.run(np.ascontiguousarray(f[30:40]), 31)
    assert np.array_equal(m2.pot_cell_runoff(), o.pot_cell_runoff())
    # switched off and on again: a fresh sum
    m2.pot_cell_runoff_output(False)
    m2.pot_cell_runoff_output(True)
    assert not m2.pot_cell_runoff().any()
