"""Free-running ensembles on the GPU: 200 seeds per configuration, batched in one context.

Kermack-McKendrick final size (the reference's tests_scientific/outbreak_size.py:60-76):
one node, fully susceptible, no births, deaths, seasonality, vaccination, migration or
heterogeneity.  Each susceptible escapes with probability exp(-beta x (total infectious
agent-days) / N), so the expected final size z solves z = 1 - exp(-R0 (z + I0 / N)) with
R0 = daily infectivity x mean infectious days; an agent infected with timer T is
infectious on T days (DiseaseState precedes Transmission within a day), so R0 is taken
as infectivity_mean x the mean infection timer of the population drawn.  The ensemble
mean of 200 seeds must be within 4 standard errors (or 0.005) of z.
"""

import numpy as np
import pytest
from scipy.optimize import brentq

from laser_polio_amd import DiseaseState_ABM, SEIR_ABM, Transmission_ABM, VitalDynamics_ABM, configs
from laser_polio_amd.propertyset import PropertySet

N_AGENTS = 100_000
I0 = 20


def km_final_size(r0, i0_frac):
    if r0 <= 1.0:
        return 0.0
    return brentq(lambda z: z - (1.0 - np.exp(-r0 * (z + i0_frac))), 1e-9, 1.0)


def _pars(r0, seed, dur=730):
    return PropertySet({
        "seed": seed, "verbose": 0, "dur": dur, "init_pop": np.array([N_AGENTS]), "init_prev": [I0],
        "init_immun": 0.0, "r0": r0, "r0_scalars": np.array([1.0]), "infectivity_mean": r0 / configs.DUR_INF_MEAN,
        "individual_heterogeneity": False, "seasonal_amplitude": 0.0, "cbr": None, "distances": np.zeros((1, 1)),
        "age_model": ("exponential", 25, 80),
        "vx_prob_ri": None, "vx_prob_sia": None, "stop_if_no_cases": False,
        "dur_exp": configs.lpd.normal(mean=3, std=1), "dur_inf": configs.lpd.gamma(shape=4.51, scale=5.32),
    })


def _build(p):
    sim = SEIR_ABM(p)
    # VitalDynamics without a birth rate: no births or deaths, it carries results.pop forward
    sim.components = [VitalDynamics_ABM, DiseaseState_ABM, Transmission_ABM]
    return sim


def test_km_final_size_closed_form():
    # z = 1 - exp(-R0 z) at I0 -> 0: R0 = 2 -> 0.7968, R0 = 4 -> 0.9802
    assert km_final_size(2.0, 0.0) == pytest.approx(0.796812, abs=1e-6)
    assert km_final_size(4.0, 0.0) == pytest.approx(0.980173, abs=1e-6)
    assert km_final_size(0.9, 1e-4) == 0.0


@pytest.mark.gpu
def test_200_seed_final_size_matches_kermack_mckendrick():
    from laser_polio_amd.batch import build_replicates_device

    r0s = (2.0, 4.0)
    plist = [_pars(r0, 10_000 * i + s) for i, r0 in enumerate(r0s) for s in range(200)]
    print(f"[ensemble] {len(plist)} simulations of {N_AGENTS} agents, 730 days", flush=True)
    batch = build_replicates_device(plist, _build)
    try:
        eng = batch.prepare()
        # the infection timers as drawn, before any agent counts down
        timers = [batch.pull_sim(200 * i, names=["infection_timer", "disease_state"])[0] for i in range(len(r0s))]
        eng.run_days(0, batch.cfg.nt, use_graph=True)
        S = np.empty_like(batch.arrays["S"])
        eng.pull("S", S)
    finally:
        batch.close()
    for i, r0 in enumerate(r0s):
        st = timers[i]["disease_state"]
        mean_t = float(timers[i]["infection_timer"][st >= 0].astype(np.float64).mean())
        r0_eff = r0 / configs.DUR_INF_MEAN * mean_t
        z = km_final_size(r0_eff, I0 / N_AGENTS)
        att = 1.0 - S[-1, 200 * i: 200 * (i + 1)].astype(np.float64) / N_AGENTS
        m, se = att.mean(), att.std(ddof=1) / np.sqrt(len(att))
        print(f"[ensemble] R0 {r0} (effective {r0_eff:.4f}): mean final size {m:.5f} +- {se:.5f}, "
              f"Kermack-McKendrick {z:.5f}", flush=True)
        assert abs(m - z) < max(4 * se, 0.005), (r0, m, se, z)
