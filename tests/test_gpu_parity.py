"""Replay parity: the HIP path against the CPU oracle, bit for bit.

Both consume the same initial population (exported by the product's host
initialisation) and the same counter-based random streams, so every results
array and every agent column must be identical after the run.
"""

import numpy as np
import pytest

from cases import CASES, make
from laser_polio_amd import engine as engmod
from oracle.oracle import OracleSim

RESULT_NAMES = [
    "S", "E", "I", "R", "E_by_strain", "I_by_strain", "potentially_paralyzed", "paralyzed",
    "new_potentially_paralyzed", "new_paralyzed", "pop", "births", "deaths", "new_exposed",
    "new_exposed_by_strain", "ri_vaccinated", "ri_protected", "ri_new_exposed_by_strain", "ipv_vaccinated",
    "sia_vaccinated", "sia_protected", "sia_new_exposed_by_strain",
]


def _compare(sim, ref):
    bad = []
    for name in RESULT_NAMES:
        if sim.results.has(name):
            got = getattr(sim.results, name)
            if not np.array_equal(got, ref.res[name]):
                d = np.argwhere(got != ref.res[name])
                bad.append(f"results.{name}: {len(d)} cells differ, first {d[0].tolist()} got {got[tuple(d[0])]} want {ref.res[name][tuple(d[0])]}")
    for name in ref.cols:
        if sim.people.has(name):
            got = getattr(sim.people, name)
            want = ref.cols[name]
            if not np.array_equal(got, want):
                d = np.nonzero(got != want)[0]
                bad.append(f"people.{name}: {len(d)} slots differ, first {d[0]} got {got[d[0]]} want {want[d[0]]}")
    if not np.array_equal(sim.people.live_end, ref.live_end):
        bad.append("live_end differs")
    assert not bad, "\n".join(bad)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("use_graph", [True, False])
def test_fused_run_matches_oracle(case, use_graph):
    print(f"[parity] case={case} graph={use_graph}", flush=True)
    sim = make(case)
    ref = OracleSim(sim.export_tables(), pop0=sim.pars.init_pop).run()
    sim.run(use_graph=use_graph)
    assert sim.t == ref.ticks_run
    _compare(sim, ref)


@pytest.mark.gpu
def test_many_nodes_match_oracle():
    """8200 nodes: node lookups past the LDS-staged tables, round 1 planned by k_plan."""
    print("[parity] case=many graph=True", flush=True)
    sim = make("many")
    ref = OracleSim(sim.export_tables(), pop0=sim.pars.init_pop).run()
    assert ref.res["new_exposed"].sum() > 0
    sim.run(use_graph=True)
    assert sim.t == ref.ticks_run
    _compare(sim, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["small3", "sia"])
def test_per_component_steps_match_oracle(case):
    """Stepping components one at a time (unfused entry points) gives the same day."""
    print(f"[parity-stages] case={case}", flush=True)
    sim = make(case)
    ref = OracleSim(sim.export_tables(), pop0=sim.pars.init_pop).run()
    sim._run_per_tick()
    _compare(sim, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["campaigns", "c0mini"])
def test_one_day_calls_with_syncs_match_oracle(case):
    """One lp_run_days call per day with a column sync after each: every sync must write
    the records the day changed back into the columns, and the next call must continue
    from the device-held lists and carried statistics."""
    print(f"[parity-day-calls] case={case}", flush=True)
    sim = make(case)
    ref = OracleSim(sim.export_tables(), pop0=sim.pars.init_pop).run()
    eng = sim._ensure_engine()
    eng.run_days(0, 1, use_graph=True)
    for t in range(1, sim.nt):
        eng.run_days(t, t + 1, use_graph=True)
        eng.sync()
    sim._mark_device_newer()
    sim.t = sim.nt
    _compare(sim, ref)


def _host_edit(state, inf_timer, risk, slots):
    """The edit a host would make between days: some susceptibles become infectious."""
    sus = slots[state[slots] == 0][:40]
    state[sus] = 2
    inf_timer[sus] = 6
    risk[slots[:25]] *= 2.0


@pytest.mark.gpu
@pytest.mark.parametrize("use_graph", [True, False])
def test_host_edit_between_days_matches_oracle(use_graph):
    """Agent columns written by the host mid-run (lp_invalidate -> recount of the carried
    susceptible statistics) give the same continuation as the oracle with the same edit."""
    print(f"[parity-host-edit] graph={use_graph}", flush=True)
    sim = make("campaigns")
    k = 16
    ref = OracleSim(sim.export_tables(), pop0=sim.pars.init_pop)
    for t in range(0, k):
        ref.day(t)
    slots = np.arange(sim.people.seg_begin[1], sim.people.seg_begin[1] + 3000)
    _host_edit(ref.cols["disease_state"], ref.cols["infection_timer"], ref.cols["acq_risk_multiplier"], slots)
    for t in range(k, sim.nt):
        ref.day(t)

    eng = sim._ensure_engine()
    eng.run_days(0, k, use_graph=use_graph)
    eng.sync()
    sim._mark_device_newer()
    p = sim.people
    _host_edit(p.disease_state, p.infection_timer, p.acq_risk_multiplier, slots)
    eng = sim._ensure_engine()          # pushes the edited columns (and invalidates)
    eng.run_days(k, sim.nt, use_graph=use_graph)
    eng.sync()
    sim._mark_device_newer()
    _compare(sim, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("case,k", [("campaigns", 7), ("campaigns", 10), ("campaigns", 14), ("sia", 14),
                                    ("small3", 9), ("ri_only", 15)])
def test_stage_calls_between_fused_runs_match_oracle(case, k):
    """A day run as stage calls between two fused runs (the path of a seed_schedule day)
    hands the device-held state (E/I lists, RI candidates, reference-tick ri_timer,
    agent records) from the fused path to the per-stage path and back."""
    print(f"[parity-mixed] case={case} k={k}", flush=True)
    sim = make(case)
    ref = OracleSim(sim.export_tables(), pop0=sim.pars.init_pop).run()
    eng = sim._ensure_engine()
    eng.run_days(0, k, use_graph=True)
    eng.step(k, engmod.STAGE_VITAL | engmod.STAGE_DISEASE)
    eng.step(k, engmod.STAGE_RI | engmod.STAGE_SIA | engmod.STAGE_TRANSMISSION | engmod.STAGE_LOG)
    eng.run_days(k + 1, sim.nt, use_graph=True)
    eng.sync()
    sim._mark_device_newer()
    _compare(sim, ref)
