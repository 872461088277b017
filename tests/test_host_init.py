"""Host initialisation pinned by the reference's own checks.

  * heterogeneity copula (populate_heterogeneous_values, model.py:816-866): the
    checks of tests/test_transmission.py:553-597 (mean, skew, KS lognormality,
    variance, Spearman correlation 0.8, infectivity mean and skew), for the
    reference model and for the lognormal model of the synthetic workloads;
  * gravity network: the reference's own restatement in log space,
    scripts/sandbox/debug_negative_network.py:46-75 (3 nodes);
  * row_normalizer: scripts/sandbox/debug_row_normalizer.py:28-33;
  * disease timers (model.py:568-587): truncation toward zero, clip to [0, 127],
    paralysis timer clipped to [0, infection timer].
"""

import numpy as np
import pytest
from scipy import stats

import laser_polio_amd as lp
from laser_polio_amd.model import populate_heterogeneous_values, timer_from_samples
from laser_polio_amd.network import gravity, row_normalizer


def _draw(n=400_000, **over):
    pars = lp.PropertySet(dict(lp.default_pars))
    pars <<= over
    rng = np.random.default_rng(2024)
    lp.distributions.set_seed(5)
    risk = np.zeros(n, np.float32)
    inf = np.zeros(n, np.float32)
    populate_heterogeneous_values(0, n, risk, inf, pars, rng)
    return pars, risk.astype(np.float64), inf.astype(np.float64)


def test_reference_heterogeneity_matches_reference_checks():
    """tests/test_transmission.py:553-597: risk ~ lognormal(mean 1, variance risk_mult_var 4),
    infectivity ~ Gamma(1, r0 / mean(dur_inf)), Spearman correlation corr_risk_inf = 0.8."""
    pars, risk, inf = _draw(r0=14, dur_inf=lp.constant(value=24))
    assert np.isclose(risk.mean(), 1.0, rtol=0.1)
    assert risk.mean() > np.median(risk)
    lr = np.log(risk[risk > 0])
    sub = lr[::8]
    assert stats.kstest(sub, "norm", args=(lr.mean(), lr.std())).pvalue > 0.01
    assert np.isclose(risk.var(), 4.0, rtol=0.3)
    # the normal correlation 2 sin(pi rho / 6) gives Spearman's rho exactly
    assert abs(stats.spearmanr(risk[::4], inf[::4]).correlation - 0.8) < 0.01
    assert np.isclose(inf.mean(), 14 / 24, rtol=0.02)
    assert inf.mean() > np.median(inf)
    # Gamma(shape 1) = exponential with the infectivity mean as scale
    assert stats.kstest(inf[::8], "expon", args=(0.0, 14 / 24)).pvalue > 0.01


def test_lognormal_heterogeneity_of_the_synthetic_workloads():
    """Both multipliers mean-1 lognormals with standard deviation 0.5 (infectivity scaled to
    its mean), Spearman correlation 0.8."""
    mean_inf = 14 / (4.51 * 5.32)
    pars, risk, inf = _draw(heterogeneity_model="lognormal", heterogeneity_sigma=0.5, infectivity_mean=mean_inf,
                            corr_risk_inf=0.8)
    assert np.isclose(risk.mean(), 1.0, rtol=0.01)
    assert np.isclose(risk.std(), 0.5, rtol=0.02)
    assert np.isclose(inf.mean(), mean_inf, rtol=0.01)
    assert np.isclose(inf.std() / inf.mean(), 0.5, rtol=0.02)
    s2 = np.log(1 + 0.25)
    for x in (risk, inf / mean_inf):
        assert stats.kstest(np.log(x[::8]), "norm", args=(-s2 / 2, np.sqrt(s2))).pvalue > 0.01
    assert abs(stats.spearmanr(risk[::4], inf[::4]).correlation - 0.8) < 0.01


def test_no_heterogeneity_gives_constants():
    pars, risk, inf = _draw(n=1000, individual_heterogeneity=False, r0=14, dur_inf=lp.constant(value=24))
    assert (risk == 1.0).all()
    assert np.allclose(inf, 14 / 24)


def test_gravity_matches_log_space_restatement():
    """scripts/sandbox/debug_negative_network.py:46-75: gravity == exp(log k + a log p_i +
    b log p_j - c log d_ij) off the diagonal, 0 on it, for the 3-node example."""
    pops = np.array([99510, 595855, 263884]) / 1e3
    dist = np.array([[0, 4, 66], [4, 0, 827], [66, 827, 0]])
    k, a, b, c = 3000, 1, 1, 2.0
    safe = np.where(dist == 0, 1e-6, dist)
    log_flow = np.log(k) + a * np.log(pops[:, None]) + b * np.log(pops[None, :]) - c * np.log(safe)
    want = np.exp(log_flow)
    np.fill_diagonal(want, 0)
    assert np.allclose(gravity(pops, dist, k, a, b, c), want)
    # and the entry the sandbox checks by hand
    assert np.isclose(gravity(pops, dist, k, a, b, c)[0, 2], k * (pops[0] ** a * pops[2] ** b) / dist[0, 2] ** c)


def test_row_normalizer_scales_rows_above_the_cap():
    """scripts/sandbox/debug_row_normalizer.py:28-33 on the float-promoted example."""
    net = np.array([[0, 6, 2], [10, 0, 13], [15, 10, 0]]).astype(float)
    got = row_normalizer(net, 0.3)
    rs = net.sum(axis=1)
    want = net * (0.3 / rs)[:, None]
    assert np.allclose(got, want)
    assert np.allclose(got.sum(axis=1), 0.3)
    small = np.array([[0, 0.05, 0.1], [0.01, 0, 0.02], [0.2, 0.05, 0]])
    assert np.array_equal(row_normalizer(small, 0.3), small)   # rows at or below the cap untouched


def test_timers_truncate_and_clip():
    x = np.array([-3.7, -0.2, 0.0, 0.9, 1.0, 2.99, 126.5, 127.0, 127.9, 400.0])
    assert timer_from_samples(x).tolist() == [0, 0, 0, 0, 1, 2, 126, 127, 127, 127]


def test_initial_timers_follow_their_definition():
    """model.py:568-587 on a constructed sim: exposure and infection timers are the clipped
    integer parts of their duration draws; paralysis timers lie in [0, infection timer]."""
    sim = lp.SEIR_ABM(lp.PropertySet({"seed": 3, "verbose": 0, "dur": 5, "init_pop": np.array([20000, 10000]),
                                      "dur_exp": lp.normal(mean=3, std=1), "dur_inf": lp.gamma(shape=4.51, scale=5.32)}))
    sim.components = [lp.DiseaseState_ABM, lp.Transmission_ABM]
    p = sim.people
    et, it, pt = (np.asarray(p.raw(n), np.int64) for n in ("exposure_timer", "infection_timer", "paralysis_timer"))
    assert et.max() <= 127 and it.max() <= 127
    assert (pt >= 0).all() and (pt <= it).all()
    assert abs(et.mean() - 2.5) < 0.1          # E[trunc(N(3, 1)) clipped at 0] ~ 2.5
    assert abs(it.mean() - (4.51 * 5.32 - 0.5)) < 0.5
