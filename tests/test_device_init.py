"""Population initialisation on the GPU (lp_init_population) against the host generator.

The device draws the same distributions from Philox streams, so the two populations
are compared by distribution: two-sample Kolmogorov-Smirnov on ages, ages at death,
multipliers and timer values, binomial bounds on the recovered-by-age fractions,
Spearman correlation of the copula; and exactly where the definition is exact (unborn
slots, node ids, RI timers at a fixed age, initial infections per node).  A
device-initialised run then replays bit for bit on the oracle.
"""

import time

import numpy as np
import pytest
from scipy import stats

from backends import oracle_of
from laser_polio_amd import configs
from test_gpu_parity import _compare


def _build(device, scale=0.01, dur=40):
    pars = configs.nigeria_national_pars(scale=scale, dur=dur)
    pars.init_on_device = device
    sim = configs.build(pars)
    if device:
        sim._ensure_engine()
        sim._pull_all()
    return sim


def _initial(sim):
    p = sim.people
    return np.concatenate([np.arange(p.seg_begin[n], p.live_end[n]) for n in range(len(p.seg_begin))])


def _unborn(sim):
    p = sim.people
    return np.concatenate([np.arange(p.live_end[n], p.seg_end[n]) for n in range(len(p.seg_begin))])


@pytest.mark.gpu
def test_device_population_matches_host_distributions():
    print("[device-init] national geography, scale 0.01", flush=True)
    host, dev = _build(False), _build(True)
    ih, idv = _initial(host), _initial(dev)
    assert np.array_equal(ih, idv)
    rng = np.random.default_rng(0)
    sub = rng.choice(len(ih), size=min(200_000, len(ih)), replace=False)
    ph, pd = host.people, dev.people

    def ks(a, b):
        return stats.ks_2samp(np.asarray(a, np.float64), np.asarray(b, np.float64)).pvalue

    age_h, age_d = -ph.date_of_birth[ih].astype(np.int64), -pd.date_of_birth[idv].astype(np.int64)
    assert ks(age_h[sub], age_d[sub]) > 1e-3
    assert age_d.min() >= 1 and age_d.max() <= 80 * 365
    aad_h = (ph.date_of_death[ih] - ph.date_of_birth[ih]).astype(np.int64)
    aad_d = (pd.date_of_death[idv] - pd.date_of_birth[idv]).astype(np.int64)
    assert (aad_d > age_d).all()                                  # death after the current age
    assert ks(aad_h[sub], aad_d[sub]) > 1e-3
    # recovered: everyone over 15 y, 80% of the under-5s, nobody in between (init_immun 0)
    st_h, st_d = ph.disease_state[ih], pd.disease_state[idv]
    for lo, hi in ((0, 5 * 365), (5 * 365, 15 * 365), (15 * 365, 200 * 365)):
        mh, md = (age_h >= lo) & (age_h < hi), (age_d >= lo) & (age_d < hi)
        fh, fd = np.mean(st_h[mh] == 3), np.mean(st_d[md] == 3)
        se = np.sqrt(fh * (1 - fh) / mh.sum() + fd * (1 - fd) / md.sum()) + 1e-12
        assert abs(fh - fd) < 5 * se + 1e-9, (lo, hi, fh, fd)
    # initial infections: the same exact count per node (chosen on the host in both paths)
    nodes = np.repeat(np.arange(len(ph.seg_begin)), ph.live_end - ph.seg_begin)
    assert np.array_equal(np.bincount(nodes[st_h == 2], minlength=len(ph.seg_begin)),
                          np.bincount(nodes[st_d == 2], minlength=len(ph.seg_begin)))
    # multipliers over every slot (newborns inherit theirs): lognormal, mean 1, sd 0.5, Spearman 0.8
    allh = rng.choice(ph.capacity, size=200_000, replace=False)
    for col in ("acq_risk_multiplier", "daily_infectivity"):
        assert ks(getattr(ph, col)[allh], getattr(pd, col)[allh]) > 1e-3, col
    rs = stats.spearmanr(pd.acq_risk_multiplier[allh], pd.daily_infectivity[allh]).correlation
    assert abs(rs - 0.8) < 0.01
    # timers: integer-valued, compared by their value frequencies
    for col in ("exposure_timer", "infection_timer", "paralysis_timer"):
        a, b = getattr(ph, col)[allh].astype(np.int64), getattr(pd, col)[allh].astype(np.int64)
        assert b.max() <= 127
        ct = np.vstack([np.bincount(a, minlength=128), np.bincount(b, minlength=128)])
        ct = ct[:, ct.sum(axis=0) >= 10]
        assert stats.chi2_contingency(ct).pvalue > 1e-3, col
    assert (pd.paralysis_timer[allh] <= pd.infection_timer[allh]).all()
    # exact parts
    assert np.array_equal(pd.node_id[idv], ph.node_id[ih])
    un = _unborn(dev)
    assert (pd.disease_state[un] == -1).all() and (pd.node_id[un] == -1).all() and (pd.date_of_birth[un] == -1).all()
    assert (pd.ri_timer[un] == -1).all() and (pd.date_of_death[un] == 0).all()
    # RI at a fixed 42 days of age: ri_timer = trunc(dob + 42)
    assert np.array_equal(pd.ri_timer[idv], np.trunc(pd.date_of_birth[idv].astype(np.float64) + 42).astype(np.int16))
    for col in ("strain", "paralyzed", "ipv_protected", "chronically_missed"):
        assert (getattr(pd, col)[idv] == 0).all(), col
    assert (pd.potentially_paralyzed[idv] == -1).all()


@pytest.mark.gpu
def test_device_initialised_run_matches_oracle():
    """The population drawn on the GPU, run 40 days, replays bit for bit on the oracle."""
    print("[device-init] replay parity", flush=True)
    sim = _build(True, scale=0.004, dur=40)
    ref = oracle_of(sim).run()
    sim.run()
    _compare(sim, ref)


@pytest.mark.gpu
def test_national_population_initialises_in_seconds():
    """The national workload (2.27e8 agents) constructed with its population drawn on the GPU."""
    t0 = time.perf_counter()
    pars = configs.nigeria_national_pars(scale=1.0)
    pars.init_on_device = True
    sim = configs.build(pars)
    t1 = time.perf_counter()
    sim._ensure_engine()
    import torch

    torch.cuda.synchronize()
    t2 = time.perf_counter()
    print(f"[device-init] national: host construction {t1 - t0:.2f}s, device draws + upload {t2 - t1:.2f}s", flush=True)
    assert t2 - t0 < 10.0
    st = sim._engine.tensors["disease_state"]
    n_live = int((st >= 0).sum())
    assert n_live == int(pars.init_pop.sum())
