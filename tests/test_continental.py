"""The continental stress case (configs.continental_pars): 5000 nodes, three strains at
relative infectivity 1 / 0.25 / 0.125, strain-specific SIA every 120 days, births and
deaths, 365 days.

  * its construction follows the configuration (CPU);
  * a reduced-scale copy (population scale 0.001: the 5000 nodes, 1.4e6 agents, the
    population drawn on the GPU) replays bit for bit on the oracle: every results row,
    agent column and live end;
  * the full case (1.36e9 agents, capacity 1.47e9 slots) runs its 365 days on one GPU and
    keeps its bookkeeping: pop = pop0 + births - deaths, S + E + I + R = pop, per-strain
    rows summing to their totals, agent columns of sampled nodes counting to the last
    day's rows; all three strains transmit.
"""

import time

import numpy as np
import pytest

from backends import oracle_of
from laser_polio_amd import configs
from test_gpu_parity import _compare


def test_continental_configuration():
    p = configs.continental_pars(scale=0.001)
    assert len(p.init_pop) == 5000
    full = configs.continental_pars(scale=1.0)
    assert 1.2e9 < int(full.init_pop.sum()) < 1.5e9
    assert np.median(full.init_pop) == pytest.approx(2e5, rel=0.1)
    days = [int((np.datetime64(e["date"]) - np.datetime64("2019-01-01")).astype(int)) for e in p.sia_schedule]
    assert days == [120, 240, 360]
    assert [e["vaccinetype"] for e in p.sia_schedule] == ["mOPV2", "nOPV2", "mOPV2"]
    assert all(len(e["nodes"]) == 1500 for e in p.sia_schedule)
    assert p.strain_r0_scalars == {0: 1.0, 1: 0.25, 2: 0.125}
    sim = configs.build(p)
    tb = sim._export_tables()
    assert tb.n_strains == 3 and list(tb.sia_strain) == [1, 2, 1]
    # replicate seeds share the geography and differ in the simulation seed
    q = configs.continental_pars(scale=0.001, run_seed=43)
    assert np.array_equal(q.init_pop, p.init_pop) and q.seed != p.seed


@pytest.mark.gpu
def test_continental_reduced_scale_matches_oracle():
    t0 = time.perf_counter()
    pars = configs.continental_pars(scale=0.001)
    pars.init_on_device = True
    sim = configs.build(pars)
    ref = oracle_of(sim)       # starts from the device-drawn columns
    print(f"[continental] scale 0.001: {int(pars.init_pop.sum())} agents; oracle start {time.perf_counter() - t0:.1f}s",
          flush=True)
    ref.run()
    print(f"[continental] oracle {ref.ticks_run} ticks {time.perf_counter() - t0:.1f}s", flush=True)
    sim.run()
    tx = (ref.res["new_exposed_by_strain"] - ref.res["sia_new_exposed_by_strain"]).sum(axis=(0, 1))
    assert (tx > 0).all(), tx
    _compare(sim, ref)


@pytest.mark.gpu
def test_continental_full_scale_keeps_its_bookkeeping():
    import torch

    import bench

    torch.cuda.empty_cache()
    print("[continental] full scale", flush=True)
    out = bench.continental_single(1.0)
    print(f"[continental] {out['agents']} agents, {out['days']} days in {out['seconds']:.3f}s; {out['properties']}",
          flush=True)
    assert out["agents"] > 1.2e9
    assert out["properties"]["ok"], out["properties"]
    assert all(x > 0 for x in out["transmission_exposures_by_strain"]), out["transmission_exposures_by_strain"]
