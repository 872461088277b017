"""Batched replicates: many independent simulations advanced by one library context.

Each simulation is an ordinary ``SEIR_ABM`` built and initialised on the host
(the reference runs replicate and calibration sims one process at a time,
calib/ and run_sim.py).  ``BatchedSims`` packs them into one set of device
buffers: node segments, node tables and network blocks are concatenated, and
every simulation keeps its own seed, slot and node origin in its
``lp_sim_params`` record.  Every kernel launch of a simulated day then covers
all simulations, so small simulations fill the GPU together.

Random draws are keyed by (sim seed, sim-local slot or node index, day), so a
simulation's trajectory in a batch is bit-identical to running it alone.

Requirements: every simulation has the same number of ticks and strains, the
same day schedule (vital-dynamics / RI step days), the stock component classes
and no ``seed_schedule`` injections (those run host code between stages).
SIA campaigns must be the same events (day, vaccine, ages) in every simulation;
which nodes they target may differ.
"""

import numpy as np

from . import engine as engmod
from .model import AGENT_COLUMNS, RESULT_ARRAYS, SEIR_ABM


class BatchedSims:
    def __init__(self, sims, tiles_per_block=0, timing=False):
        if not sims:
            raise ValueError("BatchedSims needs at least one simulation")
        self.sims = list(sims)
        for s in self.sims:
            if not isinstance(s, SEIR_ABM) or not s._fused_ok():
                raise ValueError("every simulation must be a SEIR_ABM with the stock components")
            ds = s._instance("DiseaseState_ABM")
            if ds is not None and len(ds.seed_schedule):
                raise ValueError("seed_schedule injections are not supported in a batch")
            if s.t != 0:
                raise ValueError("simulations must not have been run")
        self.tables = [s.export_tables() for s in self.sims]
        t0 = self.tables[0]
        for tb in self.tables[1:]:
            if tb.nt != t0.nt or tb.n_strains != t0.n_strains:
                raise ValueError("all simulations need the same duration and number of strains")
            if not np.array_equal(tb.day_flags, t0.day_flags):
                raise ValueError("all simulations need the same day schedule (component set, step sizes, SIA days)")
            if tb.n_sia_events != t0.n_sia_events or not (
                np.array_equal(tb.sia_strain, t0.sia_strain) and np.array_equal(tb.sia_min_age, t0.sia_min_age)
                and np.array_equal(tb.sia_max_age, t0.sia_max_age) and np.array_equal(tb.day_sia_begin, t0.day_sia_begin)
                and np.array_equal(tb.day_sia_count, t0.day_sia_count)):
                raise ValueError("SIA events must be the same in every simulation")
        self.timing = timing
        self.tiles_per_block = int(tiles_per_block)
        self._pack()
        self.engine = None

    # -- packing --------------------------------------------------------------------
    def _pack(self):
        tbs, S = self.tables, len(self.tables)
        nt, K = tbs[0].nt, tbs[0].n_strains
        caps = np.array([tb.capacity for tb in tbs], np.int64)
        nns = np.array([tb.n_nodes for tb in tbs], np.int64)
        self.slot_off = np.concatenate([[0], np.cumsum(caps)])
        self.node_off = np.concatenate([[0], np.cumsum(nns)])
        net_sizes = nns * nns
        net_off = np.concatenate([[0], np.cumsum(net_sizes)])
        N = int(self.node_off[-1])
        arrays = {name: np.concatenate([tb.cols[name] for tb in tbs]) for name in AGENT_COLUMNS}
        arrays["seg_begin"] = np.concatenate([tb.seg_begin + o for tb, o in zip(tbs, self.slot_off)])
        arrays["live_end"] = np.concatenate([tb.live_end + o for tb, o in zip(tbs, self.slot_off)])
        arrays["seg_end"] = np.concatenate([tb.seg_end + o for tb, o in zip(tbs, self.slot_off)])
        arrays["node_sim"] = np.repeat(np.arange(S, dtype=np.int32), nns)
        for key, attr in (("r0_scalar", "r0s"), ("keep", "keep"), ("birth_rate_step", "birth_rate_step"),
                          ("thr_ri", "thr_ri"), ("thr_ipv", "thr_ipv")):
            arrays[key] = np.concatenate([getattr(tb, attr) for tb in tbs])
        sp = np.concatenate([s.sim_params(tb) for s, tb in zip(self.sims, tbs)])
        sp["slot_base"] = self.slot_off[:-1]
        sp["net_off"] = net_off[:-1]
        sp["node_begin"] = self.node_off[:-1]
        arrays["sim_params"] = sp
        arrays["network_t"] = np.concatenate([tb.network_t for tb in tbs])
        arrays["km_cd"] = np.concatenate([tb.km_cd for tb in tbs])
        arrays["day_flags"] = tbs[0].day_flags
        arrays["day_sia_begin"] = tbs[0].day_sia_begin
        arrays["day_sia_count"] = tbs[0].day_sia_count
        arrays["season"] = np.ascontiguousarray(np.stack([tb.season for tb in tbs], axis=1)).reshape(-1)
        E = tbs[0].n_sia_events
        if E:
            arrays["sia_strain"] = tbs[0].sia_strain
            arrays["sia_min_age"] = tbs[0].sia_min_age
            arrays["sia_max_age"] = tbs[0].sia_max_age
            arrays["sia_thr"] = np.concatenate([tb.sia_thr.reshape(E, tb.n_nodes, 2) for tb in tbs], axis=1).reshape(-1)
        for attr, (bind, per_strain) in RESULT_ARRAYS.items():
            parts = [s._results_arrays()[bind] for s in self.sims]
            arrays[bind] = np.ascontiguousarray(np.concatenate(parts, axis=1))
        self.arrays = arrays
        self.cfg = engmod.LPConfig(capacity=int(self.slot_off[-1]), n_nodes=N, n_sims=S, n_strains=K, nt=nt,
                                   n_sia_events=E, tiles_per_block=self.tiles_per_block, mix_mode=0)

    # -- running ----------------------------------------------------------------------
    def prepare(self):
        """Upload the packed buffers (not timed by run())."""
        if self.engine is None:
            self.engine = engmod.Engine(self.cfg, self.arrays, timing=self.timing)
        return self.engine

    def run(self, use_graph=True, scatter=True):
        """Run every simulation to its end (or early stop) and, with `scatter`, copy each
        simulation's agent columns and results back into its own SEIR_ABM."""
        eng = self.prepare()
        eng.run_days(0, self.cfg.nt, use_graph=use_graph)
        eng.sync()
        self.stop_days = eng.stop_days()
        if scatter:
            self.scatter()
        return self

    def scatter(self):
        eng = self.engine
        cols = {}
        for name in AGENT_COLUMNS:
            out = np.empty_like(self.arrays[name])
            eng.pull(name, out)
            cols[name] = out
        live_end = np.empty_like(self.arrays["live_end"])
        eng.pull("live_end", live_end)
        res = {}
        for attr, (bind, _) in RESULT_ARRAYS.items():
            out = np.empty_like(self.arrays[bind])
            eng.pull(bind, out)
            res[attr] = out
        for k, sim in enumerate(self.sims):
            a, b = self.slot_off[k], self.slot_off[k + 1]
            na, nb = self.node_off[k], self.node_off[k + 1]
            for name in AGENT_COLUMNS:
                if sim.people.has(name):
                    sim.people.raw(name)[...] = cols[name][a:b]
            sim.people.live_end[...] = live_end[na:nb] - a
            for attr in RESULT_ARRAYS:
                if sim.results.has(attr):
                    sim.results.raw(attr)[...] = res[attr][:, na:nb]
            stop = int(self.stop_days[k])
            if stop >= 0:
                sim.t, sim.should_stop = stop + 1, True
            else:
                sim.t = sim.nt

    def agent_days(self, t0=1, t1=None):
        """Sum over ticks [t0, t1) of live agents of every simulation (from results.pop)."""
        pop = np.empty_like(self.arrays["pop"])
        self.engine.pull("pop", pop)
        t1 = self.cfg.nt if t1 is None else t1
        return int(pop[t0:t1].astype(np.int64).sum())

    def close(self):
        if self.engine is not None:
            self.engine.close()
            self.engine = None
