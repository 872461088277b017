"""Exact checkpoints (the resume half of the snapshot row): SEIR_ABM.save_checkpoint /
load_checkpoint.  A run stopped at day 30, saved, loaded into a freshly built model and run
on ends bit-identical to the run never stopped -- every agent column over the whole
capacity, every results array and every node's live end -- with a seed_schedule injection
after the checkpoint (host generator state restored) and RI campaigns on both sides of it.
"""

import numpy as np
import pytest

from laser_polio_amd import configs


def _pars(dur=70):
    p = configs.spatial14_pars(scale=0.02, dur=dur)
    p.vx_prob_ri = np.full(14, 0.4)
    p.ri_age_days = (42, 42)
    p.seed_schedule = [{"timestep": 45, "node_id": 3, "prevalence": 12}, {"timestep": 46, "node_id": 7, "prevalence": 0.001}]
    return p


def test_checkpoint_round_trip_before_running(tmp_path):
    a = configs.build(_pars())
    path = tmp_path / "ckpt.npz"
    a.save_checkpoint(path)
    b = configs.build(_pars())
    b.people.raw("acq_risk_multiplier")[...] = 0.0
    b.rng.random(5)
    b.load_checkpoint(path)
    for name in a.people.column_names():
        assert np.array_equal(a.people.raw(name), b.people.raw(name)), name
    assert a.rng.bit_generator.state == b.rng.bit_generator.state
    bad = configs.build(configs.spatial14_pars(scale=0.02, dur=30))
    with pytest.raises(ValueError):
        bad.load_checkpoint(path)


@pytest.mark.gpu
def test_resumed_run_equals_uninterrupted_run(tmp_path):
    print("[checkpoint] uninterrupted run", flush=True)
    a = configs.build(_pars())
    a.run()
    print("[checkpoint] run to day 30, save, load into a new model, run on", flush=True)
    b = configs.build(_pars())
    b.run(until=30)
    assert b.t == 30
    path = tmp_path / "day30.npz"
    b.save_checkpoint(path)
    c = configs.build(_pars())
    c.load_checkpoint(path)
    assert c.t == 30
    c.run()
    assert a.t == c.t == a.nt
    for name in a.people.column_names():
        assert np.array_equal(getattr(a.people, name), getattr(c.people, name)), name
    for name in a.results.column_names():
        assert np.array_equal(getattr(a.results, name), getattr(c.results, name)), name
    assert np.array_equal(a.people.live_end, c.people.live_end)
    assert a.results.new_exposed[46:].sum() > 0
