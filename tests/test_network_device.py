"""Gravity networks built on the device (lp_build_network, BatchedSims.set_gravity),
against the host construction of model.py:1213-1258 (network.gravity / row_normalizer in
float64, cast to float32) for per-simulation k and c; a batch run with the device-built
networks replays bit for bit on the oracle given the same networks.
"""

import numpy as np
import pytest

import laser_polio_amd as lp
from laser_polio_amd import configs
from laser_polio_amd.network import gravity, row_normalizer
from oracle.oracle import OracleSim


def _host_network(tb, k, a, b, c, mx):
    pop = np.asarray(tb.init_pop, np.float64)
    net = gravity(pop, tb.distances, k, a, b, c)
    net /= np.power(pop.sum(), c)
    net = row_normalizer(net, mx)
    return np.ascontiguousarray(net.T.astype(np.float32)).reshape(-1), (1.0 - net.sum(axis=1)).astype(np.float32)


def test_tables_carry_the_gravity_inputs():
    sim = configs.build(configs.replicate_pars(0, scale=0.004, dur=40))
    tb = sim.export_tables()
    k, a, b, c, mx = tb.gravity_spec
    net_t, keep = _host_network(tb, k, a, b, c, mx)
    assert np.array_equal(net_t, tb.network_t) and np.array_equal(keep, tb.keep)


@pytest.mark.gpu
def test_device_gravity_matches_host_and_replays():
    from test_gpu_parity import _compare

    print("[network] 6 replicates, per-simulation k and c", flush=True)
    sims = [configs.build(configs.replicate_pars(i, scale=0.004, dur=40)) for i in range(6)]
    tables0 = [s.export_tables() for s in sims]
    batch = lp.BatchedSims(sims)
    rng = np.random.default_rng(3)
    k = rng.uniform(0.1, 1.0, 6) * 10 ** sims[0].pars.gravity_k_exponent
    c = rng.uniform(1.5, 2.5, 6)
    specs = batch.set_gravity(k=k, c=c)
    worst = 0.0
    for i, tb in enumerate(batch.tables):
        kk, a, b, cc, mx = specs[i]
        assert kk == pytest.approx(k[i]) and cc == pytest.approx(c[i])
        net_t, keep = _host_network(tb, kk, a, b, cc, mx)
        assert np.abs(net_t).max() > 0
        # float64 on both sides, one rounding to float32: at most an ulp of float32 apart
        np.testing.assert_allclose(tb.network_t, net_t, rtol=2.5e-7, atol=1e-30)
        np.testing.assert_allclose(tb.keep, keep, rtol=2.5e-7, atol=1e-30)
        worst = max(worst, float(np.max(np.abs(tb.network_t.astype(np.float64) - net_t) / np.maximum(np.abs(net_t), 1e-30))))
    print(f"[network] worst relative difference to the host construction {worst:.2e}", flush=True)
    refs = []
    for s, tb0, tb in zip(sims, tables0, batch.tables):
        tb0.network_t, tb0.keep = tb.network_t, tb.keep
        refs.append(OracleSim(tb0, pop0=s.pars.init_pop).run())
    batch.run(use_graph=True)
    for s, ref in zip(sims, refs):
        _compare(s, ref)
    batch.close()
