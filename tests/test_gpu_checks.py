"""Checks of the HIP path beyond whole-run replay: the force of infection read back from
the device on the national geography, the float64-exact birth counts, and a caller-owned
stream handed to the library."""

import numpy as np
import pytest

import laser_polio_amd as lp
from backends import oracle_of
from cases import make
from laser_polio_amd import configs
from oracle import oracle as orc
from oracle.fp64 import foi_fp64
from test_gpu_parity import _compare


@pytest.mark.gpu
def test_foi_on_national_geography_matches_oracle_and_fp64():
    """The device's node map (Transmission_ABM.step Part 2b, model.py:1327-1363) on the 774
    nodes of the national workload (population scale 0.01, 2.3e6 agents), read through
    lp_debug_buffer after each of 12 days: p, expected, beta_mixed and the drawn counts
    bit-identical to the oracle's node map, and p and expected within 1e-6 relative of the
    float64 closed form (oracle/fp64.py) evaluated on the same tallies."""
    pars = configs.nigeria_national_pars(scale=0.01, dur=40)
    sim = configs.build(pars)
    N, K = len(pars.init_pop), len(sim.pars.strain_ids)
    print(f"[foi] national geography: {N} nodes, {int(pars.init_pop.sum())} agents", flush=True)
    tx = sim._instance("Transmission_ABM")
    ref = oracle_of(sim)
    ref.record_node_map = True
    eng = sim._ensure_engine()
    r0s = np.broadcast_to(np.asarray(tx.r0_scalars, dtype=np.float64), (N,))
    season = sim.export_tables().season
    worst, checked = 0.0, 0
    for t in range(0, 13):
        ref.day(t)
        eng.run_days(t, t + 1)
        if t == 0:
            continue
        got = {"p": eng.debug("p", N * K, np.float32).reshape(N, K),
               "expected": eng.debug("expected", N, np.float32),
               "beta_mixed": eng.debug("beta_mixed", N * K, np.float32).reshape(N, K),
               "total": eng.debug("total", N, np.int64)}
        L = ref.node_map_log[t]
        for name in ("p", "expected", "beta_mixed", "total"):
            assert np.array_equal(got[name], L[name]), f"day {t}: {name} differs from the oracle's node map"
        p64, ex64, _ = foi_fp64(L["IB"], L["SW"], tx.network, r0s, season[t], L["pop"])
        for a, b in ((got["p"], p64), (got["expected"], ex64)):
            ok = b > 1e-30
            if ok.any():
                worst = max(worst, float(np.max(np.abs(a[ok].astype(np.float64) - b[ok]) / b[ok])))
                checked += int(ok.sum())
    assert checked > 1000
    assert worst < 1e-6, worst


@pytest.mark.gpu
def test_birth_counts_are_float64_exact():
    """Expected births step_size * birth_rate * pop[t-1] (model.py:1712-1716) on 3000 nodes
    with per-node CBR in (0, 80) and populations 1..50000: the device's integer part and
    Bernoulli threshold of every node on two VD step days equal the oracle's float64
    evaluation, and the whole run's births, deaths and pop rows match."""
    rng = np.random.default_rng(23)
    n = 3000
    pop = rng.integers(1, 50_000, n)
    pop[:5] = [1, 2, 3, 7, 52143]
    xy = rng.uniform(0, 500, (n, 2))
    d = np.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1))
    sim = lp.SEIR_ABM(lp.PropertySet({"seed": 4, "verbose": 0, "dur": 15, "init_pop": pop, "cbr": rng.uniform(0, 80, n),
                                      "r0": 4, "r0_scalars": np.ones(n), "init_immun": 0.5, "init_prev": 0.001,
                                      "distances": d, "migration_method": "gravity", "stop_if_no_cases": False}))
    sim.components = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.Transmission_ABM]
    print("[births] 3000 nodes", flush=True)
    ref = oracle_of(sim).run()
    brs = sim.export_tables().birth_rate_step
    L = orc.lib()
    thr = np.zeros(1, np.uint64)
    eng = sim._ensure_engine()
    for t_vd in (7, 14):
        eng.run_days(0 if t_vd == 7 else 8, t_vd + 1)
        dbg = eng.debug("birth_expected", 2 * n, np.int64).reshape(n, 2)
        pop_prev = ref.res["pop"][t_vd - 1]
        for i in range(n):
            want = L.orc_birth_expected(float(brs[i]), int(pop_prev[i]), thr.ctypes.data)
            assert dbg[i, 0] == want and dbg[i, 1] == int(thr[0]), (t_vd, i, dbg[i].tolist(), want, int(thr[0]))
    eng.run_days(15, sim.nt)
    eng.sync()
    sim._mark_device_newer()
    sim.t = sim.nt
    _compare(sim, ref)


@pytest.mark.gpu
def test_caller_stream_is_used_and_left_to_the_caller():
    """lp_set_stream with a torch stream: the library releases its own stream, queues its
    days on the caller's, does not destroy it in lp_destroy, and the run matches the oracle."""
    import torch

    sim = make("campaigns")
    ref = oracle_of(sim).run()
    stream = torch.cuda.Stream()
    eng = sim._ensure_engine()
    assert eng.owns_stream()
    eng.set_stream(stream)
    assert not eng.owns_stream()
    eng.run_days(0, sim.nt, use_graph=True)
    eng.sync()
    sim._mark_device_newer()
    sim.t = sim.nt
    _compare(sim, ref)
    sim._drop_engine()          # lp_destroy
    with torch.cuda.stream(stream):
        x = torch.arange(1 << 20, device="cuda", dtype=torch.float32)
        y = (x * 2).sum()
    stream.synchronize()
    assert float(y) == float((np.arange(1 << 20, dtype=np.float64) * 2).sum())
