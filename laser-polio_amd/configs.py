"""Synthetic workloads: the single-simulation configurations of the benchmark.

``nigeria_national(scale)`` -- 774 nodes, populations LogNormal(ln 2.5e5, 0.6)
(about 2.3e8 agents at scale 1), coordinates uniform in a 1000 x 1000 km box,
gravity network k=0.5 a=b=1 c=2 normalised by (sum pop)^c and capped at row sum
0.3, R0 14 x per-node LogNormal(0, 0.3) scalars, seasonality amplitude 0.4
peaking on day-of-year 180, exposure Normal(3, 1) d, infectious
Gamma(4.51, 5.32) d, acquisition risk and infectivity mean-1 lognormals (sd 0.5)
with copula correlation 0.8, ages Exponential(25 y) capped at 80 y, everyone
over 15 y and 80% of the under-5s recovered, 0.5% of node 0 infectious, CBR 35,
paralysis 1/2000, RI at 6 weeks with per-node coverage Beta(8, 3) x efficacy
0.5, SIA campaigns every 180 days on a random 40% of nodes at 80% coverage of
under-5s (nOPV2), 2555 days, seed 42.

``spatial14(scale)`` -- the 14-node case: populations LogNormal(ln 3e5, 0.5)
(about 4.5e6 agents), symmetric distances U(20, 250) km, no RI or SIA, 365 days;
otherwise as above.

``continental_pars(scale, seed, run_seed)`` -- the continental stress case: 5000
nodes, populations LogNormal(ln 2e5, 0.8) (about 1.4e9 agents at scale 1),
coordinates uniform in a 4000 x 4000 km box, three strains at relative
infectivity 1 / 0.25 / 0.125 (VDPV2, Sabin2, nOPV2), strain-specific SIA
campaigns every 120 days on a random 30% of nodes at 80% coverage of under-5s
(alternately mOPV2 -> Sabin2 and nOPV2 -> nOPV2), births and deaths at CBR 35,
no RI, 365 days; otherwise as the national case.  `seed` draws the geography,
`run_seed` (default `seed`) is the simulation's own seed, so replicate seeds
share one geography.

``replicate_pars(i, master_seed, scale)`` -- replicate i of the batched-replicates
workload: a spatial14-type simulation with node count uniform 8..40, total
population uniform 0.5..6 million (times `scale`), R0 uniform 8..18, seasonal
amplitude uniform 0..0.6, gravity k uniform 0.1..1 and c uniform 1.5..2.5, and its
own seed; 365 days.
"""

import numpy as np

from . import distributions as lpd
from .model import DiseaseState_ABM, RI_ABM, SEIR_ABM, SIA_ABM, Transmission_ABM, VitalDynamics_ABM
from .propertyset import PropertySet

DUR_INF_MEAN = 4.51 * 5.32


def _common(seed, r0, n, rng):
    return {
        "seed": seed,
        "verbose": 0,
        "r0": r0,
        "r0_scalars": rng.lognormal(0.0, 0.3, n),
        "seasonal_amplitude": 0.4,
        "seasonal_peak_doy": 180,
        "dur_exp": lpd.normal(mean=3, std=1),
        "dur_inf": lpd.gamma(shape=4.51, scale=5.32),
        "heterogeneity_model": "lognormal",
        "heterogeneity_sigma": 0.5,
        "corr_risk_inf": 0.8,
        "infectivity_mean": r0 / DUR_INF_MEAN,
        "age_model": ("exponential", 25, 80),
        "init_immun": 0.0,
        "init_recovered_by_age": [(15, 200, 1.0), (0, 5, 0.8)],
        "init_prev": [0.005] + [0.0] * (n - 1),
        "cbr": np.array([35.0]),
        "p_paralysis": 1 / 2000,
        "migration_method": "gravity",
        "gravity_k": 0.5,
        "gravity_a": 1,
        "gravity_b": 1,
        "gravity_c": 2.0,
        "max_migr_frac": 0.3,
        "stop_if_no_cases": False,
    }


def spatial14_pars(scale=1.0, seed=42, dur=365):
    rng = np.random.default_rng(seed)
    n = 14
    pop = np.maximum(1, np.round(rng.lognormal(np.log(3e5 * scale), 0.5, n))).astype(np.int64)
    d = rng.uniform(20, 250, (n, n))
    d = np.triu(d, 1)
    d = d + d.T
    p = _common(seed, 14, n, rng)
    p.update(dur=dur, init_pop=pop, distances=d)
    return PropertySet(p)


def nigeria_national_pars(scale=1.0, seed=42, dur=2555, n=774):
    rng = np.random.default_rng(seed)
    pop = np.maximum(1, np.round(rng.lognormal(np.log(2.5e5 * scale), 0.6, n))).astype(np.int64)
    xy = rng.uniform(0, 1000, (n, 2))
    d = np.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1))
    p = _common(seed, 14, n, rng)
    cov = rng.beta(8, 3, n)
    sia = []
    for t in range(180, dur + 1, 180):
        nodes = np.sort(rng.choice(n, size=int(round(0.4 * n)), replace=False))
        day = np.datetime64("2019-01-01") + np.timedelta64(t, "D")
        sia.append({"date": str(day), "nodes": nodes.tolist(), "age_range": (0, 5 * 365 - 1), "vaccinetype": "nOPV2"})
    p.update(
        dur=dur,
        init_pop=pop,
        distances=d,
        vx_prob_ri=cov * 0.5,
        vx_prob_ipv=None,
        ri_age_days=(42, 42),
        ri_newborn_timer=42,
        vx_prob_sia=np.full(n, 0.8),
        sia_schedule=sia,
    )
    return PropertySet(p)


def continental_pars(scale=1.0, seed=42, dur=365, n=5000, run_seed=None):
    rng = np.random.default_rng(seed)
    pop = np.maximum(1, np.round(rng.lognormal(np.log(2e5 * scale), 0.8, n))).astype(np.int64)
    xy = rng.uniform(0, 4000, (n, 2))
    d = np.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1))
    p = _common(seed, 14, n, rng)
    sia = []
    vaccines = ("mOPV2", "nOPV2")
    for j, t in enumerate(range(120, dur + 1, 120)):
        nodes = np.sort(rng.choice(n, size=int(round(0.3 * n)), replace=False))
        day = np.datetime64("2019-01-01") + np.timedelta64(t, "D")
        sia.append({"date": str(day), "nodes": nodes.tolist(), "age_range": (0, 5 * 365 - 1),
                    "vaccinetype": vaccines[j % 2]})
    p.update(
        seed=seed if run_seed is None else int(run_seed),
        dur=dur,
        init_pop=pop,
        distances=d,
        vx_prob_ri=None,
        vx_prob_sia=np.full(n, 0.8),
        sia_schedule=sia,
        strain_ids={"VDPV2": 0, "Sabin2": 1, "nOPV2": 2},
        strain_r0_scalars={0: 1.0, 1: 0.25, 2: 0.125},
    )
    return PropertySet(p)


def replicate_pars(i, master_seed=2024, scale=1.0, dur=365):
    rng = np.random.default_rng([master_seed, i])
    n = int(rng.integers(8, 41))
    total = rng.uniform(0.5e6, 6e6) * scale
    w = rng.lognormal(0.0, 0.5, n)
    pop = np.maximum(1, np.round(total * w / w.sum())).astype(np.int64)
    d = rng.uniform(20, 250, (n, n))
    d = np.triu(d, 1)
    d = d + d.T
    r0 = float(rng.uniform(8, 18))
    p = _common(int(rng.integers(1, 2**31)), r0, n, rng)
    p.update(dur=dur, init_pop=pop, distances=d, seasonal_amplitude=float(rng.uniform(0.0, 0.6)),
             gravity_k=float(rng.uniform(0.1, 1.0)), gravity_c=float(rng.uniform(1.5, 2.5)))
    return PropertySet(p)


def build(pars, components=None):
    """SEIR_ABM with the benchmark component list (RI/SIA only when configured)."""
    sim = SEIR_ABM(pars)
    if components is None:
        components = [VitalDynamics_ABM, DiseaseState_ABM, Transmission_ABM]
        if sim.pars.vx_prob_ri is not None:
            components.append(RI_ABM)
        if sim.pars.sia_schedule:
            components.append(SIA_ABM)
    sim.components = components
    return sim


def agent_days(results_pop, t0, t1):
    """Sum over ticks t0..t1-1 of live agents (results.pop rows)."""
    return int(np.asarray(results_pop[t0:t1], dtype=np.int64).sum())
