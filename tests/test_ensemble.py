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
from scipy import stats
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


# -- the reference's own random streams and number formats ---------------------------------------
def _six_nodes(seed, dur=120):
    rng = np.random.default_rng(2024)          # one geography for every seed
    n = 6
    pop = np.round(rng.lognormal(np.log(1e4), 0.3, n)).astype(np.int64)
    d = rng.uniform(20, 150, (n, n))
    d = np.triu(d, 1)
    d = d + d.T
    p = configs.spatial14_pars(scale=0.01, seed=seed, dur=dur)
    p.init_pop, p.distances = pop, d
    p.r0_scalars = rng.lognormal(0.0, 0.3, n)
    p.init_prev = [0.002] * n
    p.cbr = None
    p.seed = seed
    return p


def _six_build(p):
    sim = SEIR_ABM(p)
    sim.components = [VitalDynamics_ABM, DiseaseState_ABM, Transmission_ABM]
    return sim


def _refsim_totals(args):
    from oracle.refsim import ReferenceSim

    tables, pop0, seed = args
    r = ReferenceSim(tables, pop0, seed).run()
    return r.res["new_exposed"][1:].sum(axis=0)


def _compare_ensembles(a, b, label):
    """Per-node totals [seeds, nodes] of two ensembles: means within 3 standard errors per
    node and KS p > 0.01 per node (6 nodes), the all-node total within 2 standard errors."""
    out = []
    for n in range(a.shape[1]):
        se = np.sqrt(a[:, n].var(ddof=1) / len(a) + b[:, n].var(ddof=1) / len(b))
        p = stats.ks_2samp(a[:, n], b[:, n]).pvalue
        out.append((n, a[:, n].mean(), b[:, n].mean(), se, p))
        assert abs(a[:, n].mean() - b[:, n].mean()) < 3 * se, (label, out[-1])
        assert p > 0.01, (label, out[-1])
    ta, tb_ = a.sum(axis=1), b.sum(axis=1)
    se = np.sqrt(ta.var(ddof=1) / len(ta) + tb_.var(ddof=1) / len(tb_))
    assert abs(ta.mean() - tb_.mean()) < 2 * se, (label, ta.mean(), tb_.mean(), se)
    return out


def test_reference_semantics_match_the_oracle_in_distribution():
    """The restatement with the reference's own streams against the oracle (the GPU's
    draws) over 60 seeds from the same initial populations: the substitutions leave the
    per-node incidence distribution unchanged (CPU check of the checker)."""
    from oracle.oracle import OracleSim
    from oracle.refsim import ReferenceSim

    a, b = [], []
    for s in range(60):
        sim = _six_build(_six_nodes(100 + s))
        tb = sim.export_tables()
        a.append(OracleSim(tb, pop0=sim.pars.init_pop).run().res["new_exposed"][1:].sum(axis=0))
        b.append(ReferenceSim(tb, sim.pars.init_pop, 5000 + s).run().res["new_exposed"][1:].sum(axis=0))
    _compare_ensembles(np.array(a), np.array(b), "oracle vs reference streams")


@pytest.mark.gpu
def test_200_seed_ensemble_matches_reference_streams():
    """200 seeds of a 6-node epidemic on the GPU (batched) against 200 runs of the
    reference's algorithm with its own random streams from the same initial populations:
    per-node incidence means within 3 SE, KS p > 0.01, total within 2 SE."""
    import laser_polio_amd as lp

    sims = [_six_build(_six_nodes(100 + s)) for s in range(200)]
    # the reference-stream runs first: the exported tables share the agent columns, which
    # the batch writes back after its run
    print("[ensemble] 200 runs with the reference's streams on the host", flush=True)
    ref = np.array([_refsim_totals((s.export_tables(), s.pars.init_pop, 5000 + i)) for i, s in enumerate(sims)])
    print("[ensemble] 200 x 6-node simulations on the GPU", flush=True)
    b = lp.BatchedSims(sims).run(use_graph=True)
    b.close()
    gpu = np.array([np.asarray(s.results.new_exposed)[1:].sum(axis=0) for s in sims])
    rows = _compare_ensembles(gpu, ref, "GPU vs reference streams")
    for n, ma, mb, se, p in rows:
        print(f"[ensemble] node {n}: GPU {ma:.1f} reference streams {mb:.1f} (SE {se:.1f}) KS p {p:.3f}", flush=True)
