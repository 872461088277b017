"""Batched replicates: one library context advancing several simulations.

CPU tests check the packing; the GPU test checks that every simulation of a
batch ends bit-identical to the oracle run of that simulation alone (draws are
keyed by sim seed and sim-local slot / node, so the batch must not change them).
"""

import numpy as np
import pytest

import laser_polio_amd as lp
from laser_polio_amd import configs
from oracle.oracle import OracleSim


def _replicates(n, scale=0.004, dur=40):
    return [configs.build(configs.replicate_pars(i, scale=scale, dur=dur)) for i in range(n)]


def test_replicate_workload_varies_per_sim():
    ps = [configs.replicate_pars(i) for i in range(12)]
    nodes = [len(p.init_pop) for p in ps]
    assert min(nodes) >= 8 and max(nodes) <= 40 and len(set(nodes)) > 3
    tot = [p.init_pop.sum() for p in ps]
    assert 0.5e6 * 0.99 <= min(tot) and max(tot) <= 6e6 * 1.01
    assert len({p.seed for p in ps}) == 12
    assert all(8 <= p.r0 <= 18 for p in ps)


def test_batch_packing():
    sims = _replicates(3)
    b = lp.BatchedSims(sims)
    a, tb = b.arrays, b.tables
    caps = [t.capacity for t in tb]
    assert b.cfg.capacity == sum(caps) and b.cfg.n_sims == 3
    assert b.cfg.n_nodes == sum(t.n_nodes for t in tb)
    sp = a["sim_params"]
    assert list(sp["slot_base"]) == [0, caps[0], caps[0] + caps[1]]
    assert list(sp["node_begin"]) == list(np.cumsum([0] + [t.n_nodes for t in tb])[:-1])
    assert list(sp["net_off"]) == list(np.cumsum([0] + [t.n_nodes ** 2 for t in tb])[:-1])
    assert list(sp["seed"]) == [t.seed for t in tb]
    for k, t in enumerate(tb):
        na, nb = b.node_off[k], b.node_off[k + 1]
        assert np.array_equal(a["seg_begin"][na:nb], t.seg_begin + b.slot_off[k])
        assert np.array_equal(a["node_sim"][na:nb], np.full(t.n_nodes, k))
        assert np.array_equal(a["season"].reshape(t.nt, 3)[:, k], t.season)
        assert np.array_equal(a["disease_state"][b.slot_off[k]:b.slot_off[k + 1]], t.cols["disease_state"])
        assert np.array_equal(a["pop"][0, na:nb], sims[k].pars.init_pop)


def _same_geography(n, i=0, scale=0.004, dur=40):
    """`n` seeds of replicate `i`'s geography (identical networks)."""
    sims = []
    for k in range(n):
        p = configs.replicate_pars(i, scale=scale, dur=dur)
        p.seed = 7000 + k
        sims.append(configs.build(p))
    return sims


def test_shared_networks_are_stored_once():
    sims = _same_geography(3) + [configs.build(configs.replicate_pars(1, scale=0.004, dur=40))]
    tb = [s.export_tables() for s in sims]
    n0, n1 = tb[0].n_nodes, tb[3].n_nodes
    assert len({t.seed for t in tb}) == 4
    assert all(np.array_equal(t.network_t, tb[0].network_t) for t in tb[:3])
    b = lp.BatchedSims(sims, share_networks=True)
    assert list(b.net_off) == [0, 0, 0, n0 * n0]
    assert list(b.arrays["sim_params"]["net_off"]) == list(b.net_off)
    assert b.arrays["network_t"].size == n0 * n0 + n1 * n1
    for k, t in enumerate(tb):
        n = t.n_nodes
        assert np.array_equal(b.arrays["network_t"][b.net_off[k]:b.net_off[k] + n * n], t.network_t)
    with pytest.raises(ValueError, match="shares networks"):
        b.set_gravity(k=0.5)
    u = lp.BatchedSims(_same_geography(3))
    assert list(u.net_off) == [0, n0 * n0, 2 * n0 * n0]


@pytest.mark.gpu
@pytest.mark.parametrize("other", [False, True])
def test_shared_networks_match_each_sim_alone(other):
    """Seeds of one geography reading one network copy: each ends bit-identical to the
    oracle run of that simulation alone (which holds its own copy), with and without
    another geography in the batch."""
    print(f"[batch-shared-network] 3 seeds, other geography {other}", flush=True)
    from test_gpu_parity import _compare

    sims = _same_geography(3)
    if other:
        sims.append(configs.build(configs.replicate_pars(1, scale=0.004, dur=40)))
    refs = [OracleSim(s.export_tables(), pop0=s.pars.init_pop).run() for s in sims]
    b = lp.BatchedSims(sims, share_networks=True)
    assert len(set(b.net_off.tolist())) == (2 if other else 1)
    b.run(use_graph=True)
    for s, ref in zip(sims, refs):
        assert s.t == ref.ticks_run
        _compare(s, ref)
    b.close()


def test_batch_rejects_mismatched_schedules():
    a = configs.build(configs.replicate_pars(0, scale=0.004, dur=40))
    b = configs.build(configs.replicate_pars(1, scale=0.004, dur=50))
    with pytest.raises(ValueError):
        lp.BatchedSims([a, b])


@pytest.mark.gpu
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_each_sim_alone(use_graph):
    print(f"[batch-parity] graph={use_graph}", flush=True)
    from test_gpu_parity import _compare

    sims = _replicates(5)
    refs = [OracleSim(s.export_tables(), pop0=s.pars.init_pop).run() for s in sims]
    b = lp.BatchedSims(sims).run(use_graph=use_graph)
    for s, ref in zip(sims, refs):
        assert s.t == ref.ticks_run
        _compare(s, ref)
    b.close()


def test_parallel_build_matches_serial_packing():
    from laser_polio_amd.batch import build_replicates

    pars = [configs.replicate_pars(i, scale=0.003, dur=30) for i in range(5)]
    par = build_replicates(pars, configs.build, workers=3)
    try:
        ser = lp.BatchedSims([configs.build(p) for p in pars])
        assert par.cfg.capacity == ser.cfg.capacity and par.cfg.n_nodes == ser.cfg.n_nodes
        for name, arr in ser.arrays.items():
            if name == "sim_params":
                assert par.arrays[name].tobytes() == arr.tobytes()
            else:
                assert np.array_equal(par.arrays[name], arr), name
    finally:
        par.close()


@pytest.mark.gpu
def test_device_initialised_replicates_full_size():
    """The replicate workload at full size (512 simulations of 8-40 nodes, 1.6e9 agents,
    365 days) with every population drawn on the GPU: the batch bookkeeping of every
    simulation (pop[t] = pop[t-1] + births - deaths; S + E + I + R = pop) and two sampled
    simulations replayed alone on the CPU oracle, bit for bit."""
    import time

    import torch

    from laser_polio_amd import configs
    from laser_polio_amd.batch import build_replicates_device
    import bench

    t0 = time.perf_counter()
    pars_list = [configs.replicate_pars(i, scale=1.0, dur=365) for i in range(512)]
    batch = build_replicates_device(pars_list, configs.build)
    eng = batch.prepare()
    torch.cuda.synchronize()
    print(f"[replicates] 512 simulations, capacity {batch.cfg.capacity}, built {time.perf_counter() - t0:.1f}s",
          flush=True)
    ks = [0, 341]
    init_state = {k: batch.pull_sim(k)[0] for k in ks}
    eng.run_days(0, batch.cfg.nt, use_graph=True)
    eng.sync()
    print(f"[replicates] ran {time.perf_counter() - t0:.1f}s", flush=True)
    checks = bench.batch_properties(batch, eng)
    assert checks["ok"], checks
    par = bench.replicate_parity(batch, ks, init_state, print)
    assert par["identical"], par
    batch.close()


@pytest.mark.gpu
def test_launch_shapes_do_not_change_results():
    """The workgroup counts of the list pass and the sampler (lp_config) leave every draw
    where it was: a batch with other grids still equals the oracle, and the autotuner
    picks each from its candidates."""
    print("[batch-grids] autotune and non-default grids", flush=True)
    from test_gpu_parity import _compare

    sims = _replicates(4)
    refs = [OracleSim(s.export_tables(), pop0=s.pars.init_pop).run() for s in sims]
    b = lp.BatchedSims(sims)
    best, _ = b.autotune(days=5, max_sims=4, grids=True)
    for g, cands in b.GRID_CANDIDATES.items():
        assert getattr(b.cfg, g) in cands
    b.cfg.ei_grid, b.cfg.sample_grid, b.cfg.commit_grid = 1024, 4096, 256
    b.run(use_graph=True)
    for s, ref in zip(sims, refs):
        _compare(s, ref)
    b.close()
