"""Snapshots and results files (model.py:177-195, 506-524, 1169-1188, 1529-1537,
1873-1885, 2070-2075; run_sim.py:388-431; utils.py:690-786, 1040-1160).

  * save_snapshot / load_snapshot round-trip every handed-out slot of every column,
    results.R and the parameters (r0);
  * load_sim makes every stock component with its init_from_file: loaded agents keep
    their values, slots not yet handed out get timers and multipliers drawn and -1 in
    potentially_paralyzed / paralyzed / ipv_protected, daily_infectivity is rescaled by
    r0 / old_r0, the components run in the stock order;
  * a simulation continued from a snapshot of a GPU run replays bit for bit on the oracle;
  * save_sim_results writes one row per (timestep, node) with the reference's columns and
    the temporal / regional groupings.
"""

import numpy as np
import pandas as pd
import pytest

from backends import oracle_of
from laser_polio_amd import PeopleFrame, configs, load_sim, save_sim_results
from laser_polio_amd.pars import default_run_order
from test_gpu_parity import _compare


def _pars(dur=60, r0=14):
    p = configs.spatial14_pars(scale=0.02, dur=dur)
    p.r0 = r0
    return p


def test_snapshot_round_trip(tmp_path):
    sim = configs.build(_pars())
    path = tmp_path / "init_pop.h5"
    sim.people.save_snapshot(path, sim.results.R[:], sim.pars)
    people, R, loaded = PeopleFrame.load_snapshot(path, n_ppl=sim.pars.init_pop, cbr=sim.pars.cbr, nt=sim.pars.dur + 10)
    p0 = sim.people
    assert np.array_equal(people.live_end - people.seg_begin, p0.live_end - p0.seg_begin)
    assert (people.seg_end - people.live_end >= p0.seg_end - p0.live_end - 1024).all()
    for name in p0.column_names():
        a = np.concatenate([getattr(p0, name)[b:e] for b, e in zip(p0.seg_begin, p0.live_end)])
        b = np.concatenate([getattr(people, name)[b:e] for b, e in zip(people.seg_begin, people.live_end)])
        assert a.dtype == b.dtype and np.array_equal(a, b), name
    assert np.array_equal(R, sim.results.R[:])
    assert loaded["r0"] == 14 and loaded["dur"] == sim.pars.dur


def test_load_sim_components(tmp_path):
    sim = configs.build(_pars())
    path = tmp_path / "pop.npz"
    sim.people.save_snapshot(path, sim.results.R[:], sim.pars)
    p2 = _pars(r0=7)
    s2 = load_sim(path, p2, verbose=0)
    assert [type(c).__name__ for c in s2.instances] == list(default_run_order)
    people = s2.people
    assert s2.pars.old_r0 == 14
    for n in range(3):
        a, e = int(people.seg_begin[n]), int(people.live_end[n])
        a0 = int(sim.people.seg_begin[n])
        got = people.daily_infectivity[a:e]
        want = sim.people.daily_infectivity[a0:a0 + e - a] * np.float32(7 / 14)
        assert np.array_equal(got, want)
        assert np.array_equal(people.disease_state[a:e], sim.people.disease_state[a0:a0 + e - a])
        unborn = slice(e, int(people.seg_end[n]))
        assert (people.potentially_paralyzed[unborn] == -1).all() and (people.ipv_protected[unborn] == -1).all()
        assert (people.infection_timer[unborn] > 0).mean() > 0.9
        assert (people.disease_state[unborn] == -1).all()


@pytest.mark.gpu
def test_continued_run_matches_oracle(tmp_path):
    sim = configs.build(_pars(dur=40))
    sim.run()
    path = tmp_path / "day40.npz"
    sim.people.save_snapshot(path, sim.results.R[:], sim.pars)
    s2 = load_sim(path, _pars(dur=60), verbose=0)
    ref = oracle_of(s2).run()
    s2.run()
    assert ref.res["new_exposed"].sum() > 0
    _compare(s2, ref)


def test_save_sim_results(tmp_path):
    sim = configs.build(_pars(dur=30))
    sim.pars.node_lookup = {n: {"dot_name": f"AFRO:LAND:S{n % 3}:D{n}"} for n in range(14)}
    rng = np.random.default_rng(0)
    for name in ("S", "E", "I", "R", "new_exposed", "paralyzed"):
        getattr(sim.results, name)[...] = rng.integers(0, 1000, getattr(sim.results, name).shape)
    df = save_sim_results(sim, tmp_path / "r.csv")
    assert len(df) == sim.nt * 14
    back = pd.read_csv(tmp_path / "r.csv")
    assert list(back.columns) == ["timestep", "date", "node", "dot_name", "S", "E", "I", "R", "P", "births", "deaths",
                                  "new_exposed", "potentially_paralyzed", "new_potentially_paralyzed", "new_paralyzed"]
    t, n = 17, 5
    row = back[(back.timestep == t) & (back.node == n)].iloc[0]
    assert row.S == sim.results.S[t, n] and row.P == sim.results.paralyzed[t, n] and row.dot_name == "AFRO:LAND:S2:D5"
    cfg = {"time_periods": {"bins": ["2019-01-15"], "labels": ["early", "late"]}, "grouping_level": "adm01"}
    g = save_sim_results(sim, tmp_path / "g.parquet", summary_config=cfg)
    assert set(g["time_period"].astype(str)) == {"early", "late"}
    assert (g["region"] == g["adm01"]).all() and g["adm01"].iloc[5] == "LAND:S2"
    with pytest.raises(RuntimeError):
        save_sim_results(sim, tmp_path / "r.h5")
