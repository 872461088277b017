"""init_sus_by_age (model.py:115-121, 164, 592-665, 1553-1573, 1651-1669): only the
susceptibles are agents, the immune rest of each node is carried in results.R.

Checked on the host against the definition: frame size per node = sum of n_susceptible,
ages inside the drawn bins and bin frequencies in proportion to n_susceptible (chi-square),
every agent susceptible except the initial infections, n_ipv_protected agents per bin,
R rows = (init_pop - init_sus) - cumulative pre-drawn immune deaths, day-0 deaths zero,
immune deaths in proportion to the definition's rates; and the run on both backends bit
for bit.
"""

import numpy as np
import pandas as pd
import pytest
from scipy import stats

from backends import oracle_of
from laser_polio_amd import configs
from test_gpu_parity import _compare


def _table():
    rows = []
    bins = [(0, 1), (1, 5), (5, 10), (10, 15), (15, 40)]
    rng = np.random.default_rng(7)
    for node in range(3):
        for lo, hi in bins:
            n_sus = int(rng.integers(2000, 9000)) if hi <= 15 else int(rng.integers(500, 1500))
            rows.append({"dot_name": f"AFRO:LAND:R{node}:D{node}", "node_id": node, "age_min_yr": lo, "age_max_yr": hi,
                         "n_susceptible": n_sus, "n_immune": int(rng.integers(20000, 60000)),
                         "n_ipv_protected": int(n_sus * 0.1) if hi <= 5 else 0})
    return pd.DataFrame(rows)


def _pars(dur=120):
    p = configs.spatial14_pars(scale=0.05, dur=dur)
    df = _table()
    n = 3
    p.init_pop = (df.groupby("node_id")["n_susceptible"].sum() + df.groupby("node_id")["n_immune"].sum()).to_numpy()
    p.distances = p.distances[:n, :n]
    p.r0_scalars = p.r0_scalars[:n]
    p.init_prev = [20, 0, 0]
    p.init_sus_by_age = df
    return p


def test_susceptible_only_frame_follows_the_table():
    p = _pars()
    sim = configs.build(p)
    df = p.init_sus_by_age
    people, res = sim.people, sim.results
    sus = df.groupby("node_id")["n_susceptible"].sum().to_numpy()
    assert np.array_equal(people.init_pop, sus)
    st = people.disease_state
    for n in range(3):
        sl = people.initial_slice(n)
        assert np.count_nonzero(st[sl] == 2) == (20 if n == 0 else 0)
        assert np.count_nonzero((st[sl] != 0) & (st[sl] != 2)) == 0
        rows = df[df["node_id"] == n]
        ages = -people.date_of_birth[sl].astype(np.int64)
        lo = (rows["age_min_yr"].to_numpy() * 365)
        hi = (rows["age_max_yr"].to_numpy() * 365)
        b = np.searchsorted(hi, ages, side="right")
        assert (b < len(hi)).all() and (ages >= np.maximum(lo[b], 1)).all()
        counts = np.bincount(b, minlength=len(hi))
        expect = rows["n_susceptible"].to_numpy() / rows["n_susceptible"].sum() * len(ages)
        assert stats.chisquare(counts, expect).pvalue > 1e-3
        ipv = people.ipv_protected[sl]
        for i, (_, r) in enumerate(rows.iterrows()):
            assert np.count_nonzero(ipv[b == i] == 1) == min(int(r["n_ipv_protected"]), int(counts[i]))
    immune = p.init_pop - sus
    deaths = res.deaths.astype(np.int64)
    assert (deaths[0] == 0).all()
    assert np.array_equal(res.R.astype(np.int64), immune[None, :] - np.cumsum(deaths, axis=0))
    # the pre-drawn deaths follow the definition's daily rates
    mid = (df["age_min_yr"] + df["age_max_yr"]).to_numpy() / 2
    rem = np.maximum(sim.pars.life_expectancies[df["node_id"].to_numpy()] - mid, 1.0)
    rate = np.bincount(df["node_id"].to_numpy(), weights=df["n_immune"].to_numpy() / (rem * 365), minlength=3)
    mean = deaths[1:].mean(axis=0)
    se = np.sqrt(rate / (deaths.shape[0] - 1))
    assert (np.abs(mean - rate) < 4 * se + 1e-9).all(), (mean, rate)


@pytest.mark.gpu
def test_susceptible_only_run_matches_oracle():
    sim = configs.build(_pars())
    ref = oracle_of(sim).run()
    sim.run()
    assert ref.res["new_exposed"].sum() > 0
    _compare(sim, ref)
