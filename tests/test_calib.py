"""Calibration drivers (calib/objective.py:9-80, calib/scoring.py:8-112).

CPU: the scoring functions against scipy's closed forms and the reference's rules
(nested dicts in sorted key order, clipping at 1e-6, weights, division by n), the
parameter suggestions.  GPU: a batched cross-entropy search recovers R0 of a synthetic
"observed" run; Adam on the mean-field adjoint closes the loss gap to the true parameters.
"""

import numpy as np
import pytest
from scipy import stats

from laser_polio_amd import calib, configs


def test_log_likelihood_fit_matches_closed_forms():
    actual = {"cases": [3, 0, 7], "by_region": {"b": 4, "a": 1}, "total": [10]}
    pred = {"cases": [2.5, 0.0, 9.0], "by_region": {"a": 2.0, "b": 3.0}, "total": [12.0]}
    out = calib.compute_log_likelihood_fit(actual, pred, weights={"total": 2.0})
    want_cases = -stats.poisson.logpmf([3, 0, 7], np.clip([2.5, 0.0, 9.0], 1e-6, None)).sum() / 3
    want_region = -stats.poisson.logpmf([1, 4], [2.0, 3.0]).sum() / 2
    want_total = -2.0 * stats.poisson.logpmf([10], [12.0]).sum()
    assert out["cases"] == pytest.approx(want_cases)
    assert out["by_region"] == pytest.approx(want_region)
    assert out["total"] == pytest.approx(want_total)
    assert out["total_log_likelihood"] == pytest.approx(want_cases + want_region + want_total)
    nb = calib.compute_log_likelihood_fit({"x": [5, 1]}, {"x": [4.0, 2.0]}, method="neg_binomial", dispersion=3.0,
                                          norm_by_n=False)
    assert nb["x"] == pytest.approx(-stats.nbinom.logpmf([5, 1], 3.0, 3.0 / (3.0 + np.array([4.0, 2.0]))).sum())
    with pytest.raises(ValueError):
        calib.compute_log_likelihood_fit({"x": [1]}, {"x": [1.0]}, method="gauss")


def test_compute_fit():
    a = {"x": [1.0, 4.0], "y": [2.0]}
    p = {"x": [2.0, 1.0], "y": [2.5], "z": [1.0]}
    assert calib.compute_fit(a, p) == pytest.approx(4.5)
    assert calib.compute_fit(a, p, use_squared=True, weights={"y": 4}) == pytest.approx(1 + 9 + 4 * 0.25)
    assert calib.compute_fit(a, p, normalize=True) == pytest.approx(1 / 4 + 3 / 4 + 0.5 / 2)


def test_suggest_params():
    rng = np.random.default_rng(0)
    space = {"r0": {"low": 6.0, "high": 18.0}, "n": {"low": 1, "high": 3}}
    ps = calib.suggest_params(space, rng, 200)
    assert all(6.0 <= p["r0"] <= 18.0 and isinstance(p["r0"], float) for p in ps)
    assert {p["n"] for p in ps} == {1, 2, 3}
    with pytest.raises(TypeError):
        calib.suggest_params({"x": {"low": "a", "high": "b"}}, rng, 1)


def _pars(r0, seed, scale=0.05, dur=120):
    p = configs.replicate_pars(0, scale=scale, dur=dur)
    p.r0 = float(r0)
    p.infectivity_mean = float(r0) / configs.DUR_INF_MEAN
    p.seed = int(seed)
    return p


@pytest.mark.gpu
def test_batched_search_recovers_r0():
    from laser_polio_amd.batch import build_replicates_device

    print("[calib] observed: the mean of 4 runs at R0 12", flush=True)
    obs = build_replicates_device([_pars(12.0, 777 + i) for i in range(4)], configs.build)
    eng = obs.prepare()
    eng.run_days(0, obs.cfg.nt, use_graph=True)
    runs = [calib.default_targets(obs.pull_sim(k)[1], None) for k in range(4)]
    obs.close()
    actual = {k: np.round(np.mean([r[k] for r in runs], axis=0)) for k in runs[0]}
    cal = calib.BatchedCalibration(lambda prm, rep: _pars(prm["r0"], 1000 + rep), {"r0": {"low": 6.0, "high": 18.0}},
                                   actual, n_replicates=2, seed=5)
    best = cal.run(n_generations=2, batch=16)
    gen0 = [t["score"] for t in cal.trials[:16]]
    print(f"[calib] best R0 {best['params']['r0']:.2f} score {best['score']:.1f}; generation 0 median "
          f"{np.median(gen0):.1f}", flush=True)
    assert best["score"] < np.median(gen0)
    assert abs(best["params"]["r0"] - 12.0) < 2.0


@pytest.mark.gpu
def test_meanfield_fit_closes_the_loss_gap():
    import torch

    from laser_polio_amd import adjoint as adj

    print("[calib] mean-field fit, national geography at scale 0.01", flush=True)
    sim = configs.build(configs.nigeria_national_pars(scale=0.01, dur=120))
    prob = adj.MeanFieldProblem(sim, 120, torch.float32, torch.device("cuda", 0))
    truth = prob.params()
    with torch.no_grad():
        prob.set_target(prob.incidence(truth).double().cpu().numpy())
        L_true = float(prob.loss(truth))
        start = prob.params(r0s=0.85, k=0.6, amp=1.3)
        L0 = float(prob.loss(start))
    fit, hist = calib.fit_meanfield(prob, init=start, names=("r0s", "k", "amp"), steps=150, lr=0.03)
    L1 = hist[-1]
    print(f"[calib] loss {L0:.1f} -> {L1:.1f} (truth {L_true:.1f}); k {float(start['k']):.3f} -> "
          f"{float(fit['k']):.3f} (truth {float(truth['k']):.3f}); amp {float(start['amp']):.3f} -> "
          f"{float(fit['amp']):.3f} (truth {float(truth['amp']):.3f})", flush=True)
    assert (L1 - L_true) < 0.1 * (L0 - L_true)
    assert abs(float(fit["amp"]) - float(truth["amp"])) < abs(float(start["amp"]) - float(truth["amp"]))
