"""WaterGAP input sets: the .UNF format against known bytes and numpy's big-endian
dtypes, the climate, table and parameter files, the soil and groundwater
derivations against hand-derived known answers from the reference's tables
(s_max.cpp, gw_frac.cpp), and a synthetic grid written as an input set and read
back into a model."""
import json
import os

import numpy as np
import pytest

from watergap2_amd import inputs


def test_unf_known_bytes_and_types(wg, tmp_path):
    p = str(tmp_path / "a.UNF0")
    inputs.unf_write(p, np.array([1.0, -2.5, 3.0e38], np.float32))
    assert open(p, "rb").read().hex() == "3f800000c0200000" + np.array([3.0e38], ">f4").tobytes().hex()
    p2 = str(tmp_path / "b.UNF2")
    inputs.unf_write(p2, np.array([1, -2, 300], np.int16))
    assert open(p2, "rb").read() == np.array([1, -2, 300], ">i2").tobytes()
    p4 = str(tmp_path / "c.UNF4")
    inputs.unf_write(p4, np.array([7, -1, 67420], np.int32))
    assert open(p4, "rb").read() == bytes.fromhex("00000007ffffffff" + "0001075c")
    p1 = str(tmp_path / "d.UNF1")
    inputs.unf_write(p1, np.array([5, -1, 9], np.int8))
    assert open(p1, "rb").read() == bytes([5, 255, 9])
    for path, arr in ((p, None), (p2, [1, -2, 300]), (p4, [7, -1, 67420]), (p1, [5, -1, 9])):
        got = inputs.unf_read(path, 3)
        if arr is not None:
            assert list(got) == arr
    # double grids go to float32 files and come back widened
    p5 = str(tmp_path / "e.UNF0")
    x = np.random.default_rng(0).normal(size=1000) * 1e3
    inputs.unf_write(p5, x)
    assert np.array_equal(inputs.unf_read_double(p5, 1000), x.astype(np.float32).astype(np.float64))
    assert np.array_equal(np.fromfile(p5, ">f4"), x.astype(np.float32))
    with pytest.raises(wg.WaterGapError) as e:
        inputs.unf_read(p5, 1001)  # shorter than the grid
    assert e.value.code == -5
    with pytest.raises(wg.WaterGapError):
        inputs.unf_read(str(tmp_path / "missing.UNF0"), 3)
    with pytest.raises(wg.WaterGapError):
        inputs.unf_type("G_X.UNF3")
    assert [inputs.unf_type(f"G.{k}") for k in ("UNF0", "UNF1", "UNF2", "UNF4")] == [0, 1, 2, 4]


def test_climate_year_layout(wg, tmp_path):
    """The four daily files are [N][365] float grids; the reader returns the
    library's [365][4][N] forcing layout."""
    n, year = 37, 1999
    rng = np.random.default_rng(1)
    grids = rng.normal(size=(4, n, 365)).astype(np.float32)
    for v, name in enumerate(inputs.CLIMATE_FILES):
        grids[v].astype(">f4").tofile(str(tmp_path / name.format(year)))
    f = inputs.read_climate_year(str(tmp_path), year, n)
    assert np.array_equal(f, grids.transpose(2, 0, 1))
    f2 = rng.normal(size=(365, 4, n)).astype(np.float32)
    inputs.write_climate_year(str(tmp_path), 2000, f2)
    assert np.array_equal(inputs.read_climate_year(str(tmp_path), 2000, n), f2)
    with pytest.raises(wg.WaterGapError):
        inputs.read_climate_year(str(tmp_path), 2001, n)


def test_lct_and_lai_tables(wg, tmp_path):
    lct = "\n".join(f"{i + 1} {1.0 + 0.5 * (i % 3)} {0.1 + 0.01 * i} {0.6 + 0.01 * i} {3 + i % 4} 0.{90 + i % 9}"
                    for i in range(18))
    lai = "\n".join(f"{i + 1} {0.5 + 0.25 * i} 0.{i % 10} 0.{(i * 3) % 10} {10 + 10 * (i % 3)} 0.{i % 7 + 1} 1.{i % 5}"
                    for i in range(18))
    (tmp_path / "LCT_22.DAT").write_text("# comment\n# another\n" + lct + "\n")
    (tmp_path / "LAI_22.DAT").write_text("# comment\n" + lai + "\n")
    t = inputs.read_lct_tables(str(tmp_path))
    for i in range(18):
        assert t.rooting_depth[i] == np.float32(1.0 + 0.5 * (i % 3))
        assert t.albedo[i] == float(f"{0.1 + 0.01 * i}")
        assert t.ddf[i] == 3 + i % 4 and t.emissivity[i] == float(f"0.{90 + i % 9}")
        assert t.lai[i] == np.float32(0.5 + 0.25 * i) and t.decid_frac[i] == np.float32(float(f"0.{i % 10}"))
        assert t.initial_days[i] == 10 + 10 * (i % 3) and t.kc_max[i] == float(f"1.{i % 5}")
    # a class out of order is an input error (the reference exits)
    (tmp_path / "LAI_22.DAT").write_text("# comment\n" + lai.replace("\n3 ", "\n4 ", 1) + "\n")
    with pytest.raises(wg.WaterGapError):
        inputs.read_lct_tables(str(tmp_path))
    # a written table reads back unchanged
    w = wg.World(wg.regional_spec(seed=0, nrows=2, ncols=2))
    inputs.write_lct_tables(str(tmp_path), w.lct)
    t2 = inputs.read_lct_tables(str(tmp_path))
    assert bytes(t2) == bytes(w.lct)


def test_parameter_json(wg, tmp_path):
    rng = np.random.default_rng(2)
    params = rng.uniform(0.1, 5.0, (26, 11))
    p = str(tmp_path / "params.json")
    inputs.write_parameter_json(p, params)
    assert np.array_equal(inputs.read_parameter_json(p), params)
    js = json.load(open(p))
    assert js["gammaHBV_runoff_coeff"][3] == params[0, 3] and js["precip_mult"][10] == params[25, 10]


def _lib():
    from watergap2_amd import _abi
    return _abi.load_library()


def soil_capacity(tawc, lc, m_root, rd):
    from watergap2_amd import _abi
    n = len(tawc)
    out = np.zeros(n, np.float32)
    args = [np.ascontiguousarray(tawc, np.float32), np.ascontiguousarray(lc, np.int8),
            np.ascontiguousarray(m_root, np.float64), np.ascontiguousarray(rd, np.float32)]
    rc = _lib().wg_soil_capacity(n, _abi.ptr(args[0], _abi.P_f), _abi.ptr(args[1], _abi.P_i8),
                                 _abi.ptr(args[2], _abi.P_d), _abi.ptr(args[3], _abi.P_f), _abi.ptr(out, _abi.P_f))
    assert rc == 0
    return out


def gw_factor(slope, texture, perma, aq, corr, m_gw, m_rg, time_series=1):
    from watergap2_amd import _abi
    n = len(slope)
    fg = np.zeros(n, np.float32)
    rg = np.zeros(n, np.int16)
    a = [np.ascontiguousarray(x, np.int8) for x in (slope, texture, perma, aq)]
    c = None if corr is None else np.ascontiguousarray(corr, np.float32)
    mg, mr = np.ascontiguousarray(m_gw, np.float64), np.ascontiguousarray(m_rg, np.float64)
    rc = _lib().wg_gw_factor(n, *[_abi.ptr(x, _abi.P_i8) for x in a], _abi.ptr(c, _abi.P_f), _abi.ptr(mg, _abi.P_d),
                             _abi.ptr(mr, _abi.P_d), time_series, _abi.ptr(fg, _abi.P_f), _abi.ptr(rg, _abi.P_i16))
    assert rc == 0
    return fg, rg


def test_soil_capacity_known_answers(wg):
    """Smax = TAWC x (M_ROOT_D x rooting depth of the class); TAWC < 0 (no soil
    data) gives -9999 (s_max.cpp:43-74)."""
    rd = np.full(18, 1.0, np.float32)
    rd[3] = 2.0
    rd[6] = 0.5
    smax = soil_capacity([100.0, 150.0, -9999.0, 80.0], [4, 7, 4, 1], [1.5, 1.0, 1.0, 1.25], rd)
    assert list(smax) == [300.0, 75.0, -9999.0, 100.0]


def test_groundwater_factor_known_answers(wg):
    """fg = M_GW_F x slope factor x texture factor x aquifer factor x (1 - permafrost
    and glacier share), each in [0, 1], with the tables of gw_frac.cpp:45-46, 96-108
    (slope classes 10..70 -> 1 .. 0.15, textures 10/20/30 -> 1/0.95/0.70 and Rgmax
    5/3/1.5 mm/d, interpolated); Rgmax in mm/d x 100 rounded half up."""
    cases = [
        # slope, texture, perma %, aquifer %, M_GW_F, M_RG_MAX -> fg, Rgmax x 100
        (10, 10, 0, 100, 1.0, 1.0, 1.0, 500),
        (25, 15, 20, 50, 1.0, 1.0, 0.925 * 0.975 * 0.5 * 0.8, 400),    # midway in both tables
        (0, 20, 0, 100, 1.0, 2.0, 0.95, 600),                          # slope class 0 reads as 10
        (70, 30, 0, 100, 1.0, 1.0, 0.15 * 0.70, 150),
        (40, 25, 50, 100, 1.0, 1.0, 0.75 * 0.825 * 0.5, 225),
        (10, 1, 0, 100, 1.0, 1.0, 0.0, 0),                             # texture 1: no recharge
        (10, 2, 0, 100, 1.0, 1.0, 0.95, 300),                          # texture 2: 0.95 and 3 mm/d
        (10, 0, 0, 100, 1.0, 3.0, 0.95, 300),                          # no texture: 0.95, 3 mm/d
        (10, 35, 0, 100, 1.0, 1.0, 0.0, 0),                            # beyond the table
        (10, 10, 0, 100, 2.0, 1.0, 0.95, 500),                         # fg > 1 -> 0.95
        (80, 10, 0, 100, 1.0, 1.0, 0.0, 500),                          # slope class beyond 70
    ]
    s, t, p, a, mg, mr, efg, erg = (np.array(c) for c in zip(*cases))
    fg, rg = gw_factor(s, t, p, a, None, mg, mr)
    np.testing.assert_allclose(fg, efg, rtol=3e-7, atol=0)
    assert list(rg) == list(erg)
    # the daily-WFD Rgmax table (time_series 0): 7 / 4.5 / 2.5 mm/d
    _, rg0 = gw_factor([10, 10, 10], [10, 20, 30], [0] * 3, [100] * 3, None, [1.0] * 3, [1.0] * 3, time_series=0)
    assert list(rg0) == [700, 450, 250]
    # the correction grid replaces fg by M_GW_F x correction where positive
    fgc, _ = gw_factor([10, 10], [10, 10], [0, 0], [100, 100], [0.1, 0.0], [1.0, 1.0], [1.0, 1.0])
    assert fgc[0] == np.float32(0.1) and fgc[1] == 1.0


def test_cell_geometry(wg):
    from watergap2_amd import _abi
    area = np.arange(1, 361, dtype=np.float64) * 10.0
    row = np.array([1, 180, 360], np.int16)
    cf = np.array([50.0, 100.0, 3.0])
    ll, gl, lr, rf = (np.array(v, np.float64) for v in ([1.0, 0, 2.0], [0, 0, 1.5], [0, 0, 0.5], [0, 0, 0.25]))
    ca, cout = np.zeros(3), np.zeros(3)
    P = _abi.P_d
    assert _lib().wg_cell_geometry(3, _abi.ptr(area, P), _abi.ptr(row, _abi.P_i16), _abi.ptr(cf, P), _abi.ptr(ll, P),
                                   _abi.ptr(gl, P), _abi.ptr(lr, P), _abi.ptr(rf, P), _abi.ptr(ca, P),
                                   _abi.ptr(cout, P)) == 0
    assert list(ca) == [10.0, 1800.0, 3600.0]
    assert list(cout) == [50.0, 100.0, 4.25]   # raised to the water fractions' sum


def synthetic_input_set(wg, world, tmp_path):
    """Writes the grid as a WaterGAP input set with soil and aquifer grids chosen
    per cell, and returns the directories and those grids."""
    rng = np.random.default_rng(7)
    n = world.n
    tawc = rng.uniform(20.0, 250.0, n).astype(np.float32)
    slope = rng.integers(10, 71, n).astype(np.int8)
    texture = rng.choice([0, 1, 2, 10, 15, 20, 25, 30], n).astype(np.int8)
    perma = rng.integers(0, 60, n).astype(np.int8)
    aq = rng.integers(10, 101, n).astype(np.int8)
    idir, rdir = str(tmp_path / "input"), str(tmp_path / "routing")
    inputs.write_watergap_inputs(world, idir, rdir, tawc, slope, texture, perma, aq)
    inputs.write_parameter_json(str(tmp_path / "params.json"), world.params)
    return idir, rdir, dict(tawc=tawc, slope=slope, texture=texture, perma=perma, aq=aq)


def test_input_set_round_trip(wg, oracle, tmp_path):
    """A synthetic grid written as an input set reads back into the same statics
    (doubles rounded to float32 as the files store them), with Smax, fg and Rgmax
    derived from the soil and aquifer grids, and the oracle runs on it."""
    w = wg.World(wg.regional_spec(seed=0, nrows=12, ncols=15, glolake_every=10, res_every=20))
    idir, rdir, g = synthetic_input_set(wg, w, tmp_path)
    params = inputs.read_parameter_json(str(tmp_path / "params.json"))
    iw = inputs.load_watergap_inputs(idir, rdir, params)
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64)  # noqa: E731
    assert np.array_equal(iw.texture, g["texture"])
    for k in ("row", "col", "cell_at", "aindex", "ldd", "land_cover", "elevation", "res_type",
              "start_month", "downstream", "rout_order", "built_up", "river_length", "river_slope", "roughness",
              "bankfull_flow"):
        assert np.array_equal(np.asarray(getattr(iw, k)).ravel(), np.asarray(getattr(w, k)).ravel()), k
    for k in ("cell_area", "contfreq", "glo_lake", "loc_lake", "glo_wetland", "loc_wetland", "lake_area",
              "reservoir_area", "glo_res", "stor_cap"):
        assert np.array_equal(getattr(iw, k), f32(getattr(w, k))), k
    np.testing.assert_allclose(iw.mean_outflow, w.mean_outflow, rtol=2e-7)
    assert np.array_equal(iw.params, w.params)
    rd = np.array([w.lct.rooting_depth[i] for i in range(18)], np.float32)
    assert np.array_equal(iw.smax, soil_capacity(g["tawc"], w.land_cover, w.params[3], rd))
    fg, rg = gw_factor(g["slope"], g["texture"], g["perma"], g["aq"], None, w.params[19], w.params[20])
    assert np.array_equal(iw.gw_factor, fg) and np.array_equal(iw.rgmax, rg)
    o = oracle.Oracle(iw)
    f = w.forcing(seed=0, day0=0, n_days=365)
    inputs.write_climate_year(str(tmp_path), 1990, f)
    f2 = inputs.read_climate_year(str(tmp_path), 1990, iw.n)
    assert np.array_equal(f, f2)
    o.run(f2[:30], 1)
    st = o.get_state()
    assert np.isfinite(st["soil"]).all() and np.isfinite(st["river"]).all()


@pytest.mark.gpu
def test_input_set_on_gpu(wg, oracle, tmp_path):
    """A model created from a WaterGAP input set (statics, parameters, daily climate
    files) runs on the GPU and equals the oracle on the same inputs."""
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from parity import assert_close
    print("case: input-set grid 180 cells, 60 days", flush=True)
    w = wg.World(wg.regional_spec(seed=0, nrows=12, ncols=15, glolake_every=10, res_every=20))
    idir, rdir, _ = synthetic_input_set(wg, w, tmp_path)
    iw = inputs.load_watergap_inputs(idir, rdir, inputs.read_parameter_json(str(tmp_path / "params.json")))
    inputs.write_climate_year(str(tmp_path), 1990, w.forcing(seed=0, day0=0, n_days=365))
    f = inputs.read_climate_year(str(tmp_path), 1990, iw.n)[:60]
    m = wg.WaterBalanceModel(iw)
    m.load_forcing(f)
    m.step_days(0, 1, 60)
    o = oracle.Oracle(iw)
    o.run(f, 1)
    assert_close(m.get_state(), o.get_state())
