"""What the trained surrogate gains over the search without it (GPU, small scale):
trained on basins it never sees again, surrogate-guided calibration of unseen
basins reaches the 1 % criterion in fewer model runs per basin than the
reference's search alone, and calibrates every basin the search calibrates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_needs_fewer_runs_than_the_plain_search(wg):
    from watergap2_amd import surrogate as sg
    from watergap2_amd.calibration import BasinRunner, make_basins, plain_search
    print("case: surrogate on 16 unseen basins (trained on 800 small basins)", flush=True)
    x, t = sg.generate_training_set(800, 12, seed=31, chunk=400, min_cells=30, max_cells=200)
    ok = np.isfinite(t)
    model, losses = sg.Surrogate.fit(x[ok], t[ok], epochs=(40, 20), batch=256)
    assert losses[-1] < 0.2 * losses[0]
    bs = make_basins(16, seed=777, min_cells=30, max_cells=200)
    forcing = bs.forcing(seed=11, day0=0, n_days=365 * 3)
    feats = sg.basin_features(bs, forcing)
    hidden = np.random.default_rng(3).uniform(0.3, 4.5, bs.n_basins)
    r = BasinRunner(bs, n_members=1, forcing_years=3, forcing=forcing)
    # 30 years, the period the network's training targets stand for
    measured = r.run(hidden[None, :], n_years=30)[0].astype(np.float32)
    plain = plain_search(r, measured, n_years=30)
    sur = sg.calibrate_with_surrogate(model, r, feats, measured)
    r.close()
    print(f"rounds {np.bincount(sur['round'], minlength=7)[1:].tolist()}; model runs per basin: surrogate "
          f"{sur['runs'].mean():.2f}, plain {plain['runs'].mean():.2f}", flush=True)
    assert sur["within_1pc"][plain["status"] == 1].all()
    assert sur["runs"].mean() < plain["runs"].mean()


def test_committed_weights_on_unseen_basins(wg):
    """The committed network (trained by watergap2_amd.train_surrogate on 5,000
    basins) on 60 unseen basins cut from another DEM: its first proposals land
    within 1 % for a fair share of basins, the runs per basin fall well below the
    plain search's, and every basin is calibrated."""
    import os
    from watergap2_amd import surrogate as sg
    from watergap2_amd.calibration import BasinRunner, make_basins, plain_search
    path = os.path.join(os.path.dirname(sg.__file__), "data", "surrogate_weights.npz")
    model, info = sg.Surrogate.load(path)
    print("case: committed surrogate weights on 60 unseen basins, 30 years", flush=True)
    bs = make_basins(60, seed=4242)
    forcing = bs.forcing(seed=11, day0=0, n_days=365 * 3)
    feats = sg.basin_features(bs, forcing)
    hidden = np.random.default_rng(8).uniform(0.1, 5.0, bs.n_basins)
    r = BasinRunner(bs, n_members=1, forcing_years=3, forcing=forcing)
    measured = r.run(hidden[None, :], n_years=30)[0].astype(np.float32)
    plain = plain_search(r, measured, n_years=30)
    sur = sg.calibrate_with_surrogate(model, r, feats, measured, first_spread=float(info["residual_sd"]))
    r.close()
    first = float((sur["round"] == 1).mean())
    print(f"first-proposal hit rate {first:.2f}; runs per basin: surrogate {sur['runs'].mean():.2f}, "
          f"plain {plain['runs'].mean():.2f}; median |gamma error| "
          f"{np.median(np.abs(sur['gamma'] - hidden)):.3f}", flush=True)
    assert sur["within_1pc"].all()
    assert first >= 0.08
    assert sur["runs"].mean() <= 0.6 * plain["runs"].mean()
    assert np.median(np.abs(sur["gamma"] - hidden)) < 0.1
