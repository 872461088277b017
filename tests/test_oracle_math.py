"""Pins of the oracle's building blocks against published vectors and mathematics."""

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as orc


# -- Philox4x32-10: known-answer vectors published with Random123 (Salmon et al., SC'11) --
@pytest.mark.parametrize(
    "ctr,key,want",
    [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        (
            [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
            [0xA4093822, 0x299F31D0],
            [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1],
        ),
    ],
)
def test_philox_known_answers(ctr, key, want):
    got = orc.philox(ctr, key[0] | (key[1] << 32))
    assert [int(x) for x in got] == want


def test_philox_uniformity():
    out = np.array([orc.philox([i, 7, 3, 1], 42)[0] for i in range(20000)], dtype=np.float64) / 2**32
    assert stats.kstest(out, "uniform").pvalue > 1e-3


def _ulp_err(got, want):
    got = np.asarray(got, np.float64)
    want32 = np.asarray(want, np.float64)
    spacing = np.spacing(np.abs(want32).astype(np.float32)).astype(np.float64)
    return np.abs(got - want32) / spacing


def test_expm1_exp_log1p_accuracy():
    L = orc.lib()
    xs = np.concatenate([-np.logspace(-12, 1.2, 2000), np.linspace(-17, 5, 2000)]).astype(np.float32)
    e1 = np.array([L.orc_expm1f_c(float(x)) for x in xs], np.float32)
    assert _ulp_err(e1, np.expm1(xs.astype(np.float64))).max() <= 2.0
    xe = np.linspace(-80, 20, 3000).astype(np.float32)
    ex = np.array([L.orc_expf_c(float(x)) for x in xe], np.float32)
    assert _ulp_err(ex, np.exp(xe.astype(np.float64))).max() <= 2.0
    xl = np.concatenate([np.logspace(-10, 6, 3000), [0.0, 1.0]]).astype(np.float32)
    lg = np.array([L.orc_log1pf_c(float(x)) for x in xl], np.float32)
    assert _ulp_err(lg, np.log1p(xl.astype(np.float64))).max() <= 3.0
    # the small-rate regime that makes 1 - exp(-x) useless in float32
    assert L.orc_expm1f_c(-1e-9) == pytest.approx(-1e-9, rel=1e-7)


@pytest.mark.parametrize("lam", [0.05, 0.7, 3.0, 15.9, 16.0, 40.5, 300.0])
def test_poisson_sampler_distribution(lam):
    L = orc.lib()
    n = 6000
    x = np.array([L.orc_poisson(lam, 99, node, t) for node in range(60) for t in range(n // 60)], np.float64)
    se = np.sqrt(lam / n)
    assert abs(x.mean() - lam) < 4 * se
    assert abs(x.var() - lam) < 6 * lam * np.sqrt(2.0 / n) + 4 * se
    if lam <= 16:
        ks = np.arange(0, int(lam + 8 * np.sqrt(lam) + 6))
        obs = np.array([(x == k).sum() for k in ks], np.float64)
        exp = stats.poisson.pmf(ks, lam) * n
        keep = exp > 5
        chi = (((obs - exp) ** 2) / exp)[keep].sum()
        assert stats.chi2.sf(chi, keep.sum() - 1) > 1e-3


@pytest.mark.parametrize("r,mean", [(1000, 0.5), (1000, 7.0), (3, 2.0), (1, 30.0)])
def test_negbin_sampler_distribution(r, mean):
    L = orc.lib()
    n = 6000
    q = mean / r
    x = np.array([L.orc_negbin(r, q, 7, node, t) for node in range(60) for t in range(n // 60)], np.float64)
    p = 1.0 / (1.0 + q)
    var = r * (1 - p) / p**2
    assert abs(x.mean() - mean) < 4 * np.sqrt(var / n)
    assert abs(x.var() - var) / var < 0.15


def test_km_age_at_death_follows_table():
    """Year of death ~ deaths per year of the cumulative table, conditional on the current age."""
    import laser_polio_amd as lp

    cd = lp.create_cumulative_deaths(1_000_000, 100)
    L = orc.lib()
    cdp = np.ascontiguousarray(cd, np.int64)
    for age in (0, 20 * 365 + 100):
        ay = age // 365
        ages_at_death = np.array([L.orc_km_age_at_death(cdp.ctypes.data, len(cd), 5, s, 3, age) for s in range(20000)])
        assert (ages_at_death > age).all()
        yod = ages_at_death // 365
        lo = 0 if ay == 0 else cd[ay - 1]
        probs = np.diff(np.concatenate([[lo], cd[ay:]])) / (cd[-1] - lo)
        obs = np.bincount(yod - ay, minlength=len(probs))[: len(probs)].astype(float)
        exp = probs * len(yod)
        keep = exp > 5
        chi = (((obs - exp) ** 2) / exp)[keep].sum()
        assert stats.chi2.sf(chi, keep.sum() - 1) > 1e-3


def test_probability_thresholds():
    from laser_polio_amd import engine

    t = engine.thr_from_prob64(np.array([0.0, 1.0, 0.5, 2**-32, 1e-300, -0.1]))
    assert t.tolist() == [0, 2**32, 2**31, 1, 1, 0]


def test_expected_births_match_float64_reference_expression():
    """VitalDynamics_ABM births (model.py:1712-1716): expected = step_size * birth_rate *
    pop[t-1] evaluated by numpy in float64 (step_size an int, birth_rate = cbr / (365 * 1000)
    float64, pop int32), integer part by astype(int32), Bernoulli on expected - integer.
    The oracle's integer part must be numpy's, and its threshold ceil(frac * 2^32)."""
    rng = np.random.default_rng(11)
    cbr = np.concatenate([rng.uniform(0.0, 80.0, 3000), [0.0, 35.0, 37.0, 41.0, 1e-9, 365.0, 1000.0]])
    pop = np.concatenate([rng.integers(0, 3_000_000, 3000), [1, 7, 52143, 2**31 - 1, 1000, 365000, 3]]).astype(np.int32)
    birth_rate = cbr / (365 * 1000)
    for step in (1, 7, 14):
        expected = step * birth_rate * pop
        want_int = expected.astype(np.int32)
        frac = expected - want_int
        want_thr = np.ceil(frac * 4294967296.0).astype(np.uint64)
        L = orc.lib()
        thr = np.zeros(1, np.uint64)
        brs = step * birth_rate
        for i in range(len(pop)):
            got = L.orc_birth_expected(float(brs[i]), int(pop[i]), thr.ctypes.data)
            assert got == want_int[i], (step, cbr[i], pop[i])
            assert int(thr[0]) == int(want_thr[i]), (step, cbr[i], pop[i])


def test_km_age_at_death_brute_force_on_a_tiny_table():
    """Age at death on a 4-year table where every year gets many draws: the year of death
    follows deaths per year conditioned on the current year of age (an off-by-one in the
    conditioning moves a whole year's mass), the day is uniform over the part of that
    year after the current age, and every draw is later than the current age."""
    cd = np.ascontiguousarray([100, 300, 600, 1000], np.int64)
    L = orc.lib()
    n = 40000
    for age in (0, 400, 800, 1000):
        ay = age // 365
        draws = np.array([L.orc_km_age_at_death(cd.ctypes.data, len(cd), 9, s, 5, age) for s in range(n)])
        assert (draws > age).all()
        yod = draws // 365
        lo = 0 if ay == 0 else int(cd[ay - 1])
        p = np.diff(np.concatenate([[lo], cd[ay:]])) / (cd[-1] - lo)
        obs = np.bincount(yod - ay, minlength=len(p))
        assert len(obs) == len(p), "a year outside [current year, last year] was drawn"
        exp = p * n
        assert (exp >= 5).all()
        chi = (((obs - exp) ** 2) / exp).sum()
        assert stats.chi2.sf(chi, len(p) - 1) > 1e-3, (age, obs, exp)
        for y in range(ay, len(cd)):
            d = draws[yod == y]
            lo_d, hi_d = max(y * 365, age + 1), (y + 1) * 365
            assert d.min() >= lo_d and d.max() < hi_d
            width = hi_d - lo_d
            if len(d) >= 5 * width:
                obs_d = np.bincount(d - lo_d, minlength=width)
                e = len(d) / width
                chi_d = (((obs_d - e) ** 2) / e).sum()
                assert stats.chi2.sf(chi_d, width - 1) > 1e-4, (age, y)
