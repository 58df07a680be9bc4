"""What the trained surrogate gains over the search without it (GPU, small scale):
trained on basins it never sees again, its first gamma proposal for unseen basins
lands far closer to the target discharge than the reference search's start value,
and surrogate-guided calibration needs fewer model runs than the plain search."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_beats_search_start_on_unseen_basins(wg):
    from watergap2_amd import surrogate as sg
    from watergap2_amd.calibration import BasinRunner, make_basins, plain_search, within_1pc
    print("case: surrogate on 16 unseen basins (trained on 120 small basins)", flush=True)
    x, t = sg.generate_training_set(120, 24, seed=31, chunk=60, min_cells=30, max_cells=200)
    ok = np.isfinite(t)
    model, losses = sg.Surrogate.fit(x[ok], t[ok], epochs=(40, 20), batch=256)
    assert losses[-1] < 0.2 * losses[0]
    bs = make_basins(16, seed=777, min_cells=30, max_cells=200)
    forcing = bs.forcing(seed=11, day0=0, n_days=365 * 3)
    feats = sg.basin_features(bs, forcing)
    hidden = np.random.default_rng(3).uniform(0.3, 4.5, bs.n_basins)
    years = 9
    r = BasinRunner(bs, n_members=1, forcing_years=3, forcing=forcing)
    measured = r.run(hidden[None, :], n_years=years)[0].astype(np.float32)
    # the network predicts the 30-year mean; nine years of the 3-year forcing cycle give it
    def long_term(annual):
        return sg.runoff_target(sg.thirty_year_mean(annual)[:, None], bs)

    target = long_term(measured.astype(np.float64))
    g_star = model.propose_gamma(feats, target)
    q_star = r.run(g_star[None, :], n_years=years)[0]
    q_one = r.run(np.ones((1, bs.n_basins)), n_years=years)[0]
    miss_star = np.abs(long_term(q_star) - target)
    miss_one = np.abs(long_term(q_one) - target)
    assert np.median(miss_star) < 0.5 * np.median(miss_one), (np.median(miss_star), np.median(miss_one))
    # model runs to the 1 % criterion: plain search vs proposal + plain fallback
    plain = plain_search(r, measured, n_years=years)
    hit = np.array([within_1pc(q_star[b], measured[b]) for b in range(bs.n_basins)])
    runs_sur = np.where(hit, 1, 1 + plain["runs"])
    print(f"first-proposal hits {int(hit.sum())}/{bs.n_basins}; model runs per basin: surrogate "
          f"{runs_sur.mean():.2f}, plain {plain['runs'].mean():.2f}", flush=True)
    assert (plain["status"] == 1).all()
    r.close()
