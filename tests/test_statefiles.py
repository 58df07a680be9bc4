"""Restart files: the reference's text layouts (known text), round trips, the
conversion of a state to the files' mm values and back, and a restart through
the files that continues the run."""
import numpy as np

from parity import assert_close
from watergap2_amd import statefiles


def test_wghm_state_text_layout(wg, tmp_path):
    """Header, column labels right-aligned in 33 characters after a 6-character ID
    column, then per cell the id, TWS (the sum) and the ten storages in
    %33.16e (std::scientific, precision 16; wghmStateFile.cpp:357-413)."""
    v = np.array([[1.0, 2.0, 3.0, 0.5, 0.25, 0.0, 0.0, 0.0, 10.0, 1e-20],
                  [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.5]])
    p = str(tmp_path / "s.txt")
    statefiles.write_wghm_state(p, v)
    lines = open(p).read().split("\n")
    assert lines[0] == "WGHM water storage states"
    assert lines[1] == "    ID" + "".join(f"{h:>33}" for h in ("TWS", "CANOPY", "SNOW", "SOIL", "LOCALLAKE",
                                                                 "LOCALWETLAND", "GLOBALLAKE", "GLOBALWETLAND",
                                                                 "RESERVOIR", "RIVER", "GROUNDWATER"))
    assert lines[2] == "     1" + " " * 11 + "1.6750000000000000e+01" + "".join(f"{x:33.16e}" for x in v[0])
    assert lines[3].startswith("     2" + " " * 10 + "-1.5000000000000000e+00")
    assert np.array_equal(statefiles.read_wghm_state(p, 2), v)


def test_snow_and_additional_round_trip(wg, tmp_path):
    rng = np.random.default_rng(3)
    snow = rng.uniform(0, 3000, (7, 101))
    p = str(tmp_path / "snow.txt")
    statefiles.write_snow_elevation(p, snow)
    text = open(p).read()
    assert text.startswith("Snow in elevation for land cells\n1\t") and not text.endswith("\n")
    assert np.array_equal(statefiles.read_snow_elevation(p, 7), snow)
    add = rng.normal(size=(5, 53))
    p2 = str(tmp_path / "add.txt")
    statefiles.write_additional(p2, add)
    assert open(p2).read().split("\n")[1].startswith("    ID" + f"{'G_days_since_start':>33}")
    assert np.array_equal(statefiles.read_additional(p2, 5), add)


def test_restart_through_files_continues_the_run(wg, oracle, tmp_path):
    """A state saved to the three files after 60 days and loaded into a fresh model
    continues to the same state after 30 more days as the uninterrupted run (the
    files keep 17 significant digits; storages go through mm and back)."""
    w = wg.World(wg.regional_spec(seed=5, nrows=30, ncols=40, edge_cases=True, glolake_every=20, res_every=25))
    f = w.forcing(seed=2, day0=0, n_days=90)
    a = oracle.Oracle(w)
    a.run(f[:60], 1)
    mid = a.get_state()
    statefiles.save_state(str(tmp_path), w, mid)
    # fields the reference's files do not hold come from the running model
    base = {k: mid[k] for k in ("red_river", "river_frac_next", "storage_transfer", "started")}
    st = statefiles.load_state(str(tmp_path), w, base=dict(wg.new_state(w.n), **base))
    land = mid["laf"] > 0
    for k in ("groundwater", "loc_lake", "loc_wetl", "glo_lake", "glo_wetl", "reservoir", "river", "lai_days",
              "lai_status", "lai_precsum", "k_release", "laf", "laf_prev", "red_loc_lake", "red_glo_wetl"):
        np.testing.assert_allclose(st[k], mid[k], rtol=4e-16, atol=0, err_msg=k)
    for k in ("canopy", "soil", "snow"):
        np.testing.assert_allclose(st[k][land], mid[k][land], rtol=1e-15, atol=0, err_msg=k)
    b = oracle.Oracle(w)
    b.set_state(st)
    a.run(f[60:], 61)
    b.run(f[60:], 61)
    assert_close(b.get_state(), a.get_state())


def test_malformed_state_files_are_refused(wg, tmp_path):
    """A missing file, a truncated table (fewer rows than cells), a row with fewer
    numbers than columns, an id outside 1..N and a non-numeric row all give
    WG_E_INPUT; header lines (starting with a letter) are skipped, and a table
    ends at its first empty line, as the reference's reader does."""
    import pytest
    v = np.arange(30, dtype=np.float64).reshape(3, 10) / 7.0
    p = tmp_path / "s.txt"
    statefiles.write_wghm_state(str(p), v)
    good = p.read_text().split("\n")

    def refused(text):
        q = tmp_path / "bad.txt"
        q.write_text(text)
        with pytest.raises(wg.WaterGapError) as e:
            statefiles.read_wghm_state(str(q), 3)
        assert e.value.code == -5

    with pytest.raises(wg.WaterGapError):
        statefiles.read_wghm_state(str(tmp_path / "missing.txt"), 3)
    refused("\n".join(good[:4]) + "\n")                              # two of three rows
    refused("\n".join(good[:4] + [" ".join(good[4].split()[:6])]) + "\n")  # short row
    refused("\n".join(good[:4] + ["     9" + good[4][6:]]) + "\n")  # id 9 of 3 cells
    refused("\n".join(good[:4] + ["     3 1.0 x"]) + "\n")              # not a number
    refused("\n".join(good[:3] + [""] + good[3:]))                     # empty line ends the table early
    # extra header lines anywhere before the rows are skipped
    q = tmp_path / "hdr.txt"
    q.write_text("\n".join(good[:2] + ["Comment line"] + good[2:]))
    assert np.array_equal(statefiles.read_wghm_state(str(q), 3), v)
