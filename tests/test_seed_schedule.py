"""seed_schedule injections (DiseaseState_ABM.step, model.py:759-786), both backends.

The oracle applies the seeds from its own restatement (oracle.OracleSim._seed) with a
copy of the host Generator the product draws from, so the seeded agents, their fresh
infection timers and everything after them must match bit for bit.
"""

import numpy as np
import pytest

import laser_polio_amd as lp
from backends import BACKENDS, oracle_of, run
from cases import make
from test_gpu_parity import _compare

NAMES = ["ANKA", "BAKURA", "BIRNIN_MAGAJI", "BUKKUYUM", "BUNGUDU", "GUMMI", "GUSAU", "KAURA_NAMODA",
         "MARADUN", "MARU", "SHINKAFI", "TALATA_MAFARA", "TSAFE", "ZURMI"]


def _zamfara_like(seed=5):
    """14 synthetic nodes named like the Zamfara LGAs of the reference test, 1/100 scale."""
    rng = np.random.default_rng(seed)
    n = len(NAMES)
    lookup = {i: {"dot_name": f"AFRO:NIGERIA:ZAMFARA:{NAMES[i]}", "lat": 12.0 + rng.uniform(-1, 1),
                  "lon": 6.0 + rng.uniform(-1, 1)} for i in range(n)}
    pop = rng.integers(2000, 6000, n)
    d = rng.uniform(10, 120, (n, n))
    d = np.triu(d, 1)
    d = d + d.T
    return lookup, pop, d


@pytest.mark.parametrize("backend", BACKENDS)
def test_seed_schedule(backend):
    """tests/test_diseasestate_abm.py:398-457 with synthetic nodes: r0 = 0 (no transmission),
    init_prev 0.01 in the first node only, seeds by date + dot_name on days 1 (BAKURA) and 2
    (GUMMI) and by timestep + node_id on day 3 (node 13).  Beyond the reference's "> 0"
    checks: each seeded node's I count on its seed day is exactly int(live agents *
    prevalence), since no agent leaves I within a day of being seeded."""
    lookup, pop, d = _zamfara_like()
    init_prev = [0.01] + [0.0] * (len(pop) - 1)
    seed_schedule = [
        {"date": "2019-01-02", "dot_name": "AFRO:NIGERIA:ZAMFARA:BAKURA", "prevalence": 0.1},
        {"date": "2019-01-03", "dot_name": "AFRO:NIGERIA:ZAMFARA:GUMMI", "prevalence": 0.1},
        {"timestep": 3, "node_id": 13, "prevalence": 0.05},
    ]
    sim = lp.SEIR_ABM(lp.PropertySet({
        "seed": 17, "verbose": 0, "start_date": lp.date("2019-01-01"), "dur": 30, "init_pop": pop,
        "cbr": np.array([35.0]), "r0": 0, "r0_scalars": np.ones(len(pop)), "init_immun": 0.0, "init_prev": init_prev,
        "distances": d, "node_lookup": lookup, "seed_schedule": seed_schedule, "stop_if_no_cases": False,
        "migration_method": "gravity", "vx_prob_ri": None, "vx_prob_sia": None,
    }))
    sim.components = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.Transmission_ABM]
    out = run(sim, backend)
    I = out.res["I"]
    assert np.sum(I[0, 1:]) == 0, "There should be no infections on day 0"
    for t, node, prev in ((1, 1, 0.1), (2, 5, 0.1), (3, 13, 0.05)):
        assert I[t, node] > 0, f"no infections seeded in node {node} at t={t}"
        assert I[t - 1, node] == 0
        assert I[t, node] == int(out.res["pop"][t, node] * prev)
    assert I[1, 5] == 0 and I[2, 13] == 0


def test_seed_schedule_parsing_matches_oracle():
    """The host's parsed schedule (tick -> [(node, value)]) equals the oracle's parse of the
    same raw entries; entries whose dot_name matches no node are dropped."""
    from oracle.oracle import parse_seed_schedule

    sim = make("seeds")
    extra = list(sim.pars.seed_schedule) + [{"date": "2019-01-05", "dot_name": "AFRO:NOWHERE", "prevalence": 0.5}]
    ds = sim._instance("DiseaseState_ABM")
    ref = parse_seed_schedule({"seed_schedule": extra, "node_lookup": sim.pars.node_lookup,
                               "start_date": sim.pars.start_date})
    assert {k: v for k, v in ds.seed_schedule.items()} == ref
    assert sorted(ref) == [1, 2, 7, 10, 14]


def test_stop_rule_counts_only_seeds_after_the_day():
    """model.py:788-795: on the day of a seed the seed is not "future"; a run whose
    epidemic has died out stops on that day (after seeding it) unless a later seed exists."""
    sim = make("stopseed", seed_schedule=[{"timestep": 60, "node_id": 1, "prevalence": 20}])
    o = oracle_of(sim).run()
    assert o.stop_day == 60
    assert o.res["I"][60, 1] == 20


@pytest.mark.parametrize("case", ["seeds", "stopseed"])
def test_seeded_counts_follow_the_definition(case):
    """On the oracle: each seed day's newly infectious agents in a seeded node number
    int(live * value) (float) or min(value, live) (int)."""
    sim = make(case)
    o = oracle_of(sim)
    sched = o.seed_schedule
    for t in range(o.nt):
        before = o.cols["disease_state"].copy()
        o.day(t)
        for node, value in sched.get(t, []):
            lo, hi = int(o.seg_begin[node]), int(o.live_end[node])
            newly_i = int(np.sum((o.cols["disease_state"][lo:hi] == 2) & (before[lo:hi] != 2)))
            live = int(np.sum(o.cols["disease_state"][lo:hi] >= 0))
            want = int(live * value) if isinstance(value, float) else min(int(value), live)
            # agents already I can be drawn again; E agents turning I the same day add more
            assert newly_i >= want - int(np.sum(before[lo:hi] == 2))
    if case == "stopseed":
        assert o.stop_day is not None and o.stop_day > 62
        assert o.res["I"][60].sum() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["seeds", "stopseed"])
@pytest.mark.parametrize("mode", ["graph", "nograph", "per_tick"])
def test_seeded_run_matches_oracle(case, mode):
    """Seed days on the GPU: fused runs split at each seed day into a VD + DS stage call,
    the host seeding (pull, edit, push, lp_invalidate) and a RECOUNT stage call of
    RI + SIA + TX + log; or every component stepped on its own.  Bit-exact against the
    oracle run with the same seeds."""
    print(f"[parity-seeds] case={case} mode={mode}", flush=True)
    sim = make(case)
    ref = oracle_of(sim).run()
    assert sum(len(v) for v in ref.seed_schedule.values()) > 0
    if mode == "per_tick":
        sim._run_per_tick()
    else:
        sim.run(use_graph=(mode == "graph"))
        assert sim.t == ref.ticks_run
    _compare(sim, ref)
