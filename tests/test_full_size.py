"""Replay parity at the workloads' own sizes.

  * the 14-node spatial case (4.5e6 agents, 365 days, seed 42): the whole run;
  * the national case (774 nodes, 2.27e8 agents, capacity 2.97e8): the first 12 days,
    every results row, every agent column and every node's live end.  These days
    exercise the carried susceptible statistics, the E / I event lists and nodes of more
    than 64 spans against the oracle's dense recount of every agent every day.
"""

import time

import numpy as np
import pytest

from backends import oracle_of
from laser_polio_amd import configs
from test_gpu_parity import _compare


@pytest.mark.gpu
def test_spatial14_full_size_matches_oracle():
    t0 = time.perf_counter()
    pars = configs.spatial14_pars(scale=1.0)
    sim = configs.build(pars)
    print(f"[full-size] spatial14: {int(pars.init_pop.sum())} agents, {sim.nt - 1} days "
          f"(init {time.perf_counter() - t0:.1f}s)", flush=True)
    ref = oracle_of(sim).run()
    print(f"[full-size] oracle done {time.perf_counter() - t0:.1f}s", flush=True)
    sim.run()
    assert sim.t == ref.ticks_run
    assert ref.res["new_exposed"].sum() > 100_000
    _compare(sim, ref)


@pytest.mark.gpu
def test_national_first_days_match_oracle():
    days = 12
    t0 = time.perf_counter()
    pars = configs.nigeria_national_pars(scale=1.0)
    sim = configs.build(pars)
    print(f"[full-size] national: {int(pars.init_pop.sum())} agents, capacity {sim.people.capacity} "
          f"(init {time.perf_counter() - t0:.1f}s)", flush=True)
    ref = oracle_of(sim)
    for t in range(days + 1):
        ref.day(t)
    print(f"[full-size] oracle {days} days done {time.perf_counter() - t0:.1f}s", flush=True)
    eng = sim._ensure_engine()
    eng.run_days(0, days + 1, use_graph=True)
    eng.sync()
    sim._mark_device_newer()
    spans = (np.asarray(sim.people.seg_end) - np.asarray(sim.people.seg_begin)) // (16 * 1024)
    assert spans.max() > 64
    for name in list(ref.res):
        if sim.results.has(name):
            ref.res[name][days + 1:] = np.asarray(getattr(sim.results, name))[days + 1:]
    _compare(sim, ref)
    print(f"[full-size] compared {time.perf_counter() - t0:.1f}s", flush=True)
