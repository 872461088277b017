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

``build_replicates(pars_list, workers)`` initialises the simulations of a large
batch in worker processes that write the agent columns straight into shared
memory at their batch offsets, returning a ``BatchedSims`` without per-sim
objects.  ``build_replicates_device(pars_list)`` draws every simulation's agent
columns on the GPU instead (lp_init_population, one spec per simulation, draws
keyed by sim-local slot, so a simulation in a batch equals the same simulation
initialised alone); the host only builds node tables and schedules.

With ``share_networks=True`` simulations whose transposed networks are identical
(replicate seeds of one geography) read one copy of it: the network mix streams each
row once for all of them.  ``set_gravity`` then refuses, since it builds one network
per simulation.
"""

import multiprocessing as mp
import time
from multiprocessing import shared_memory

import numpy as np

from . import engine as engmod
from .laserframe import PeopleFrame
from .model import AGENT_COLUMNS, RESULT_ARRAYS, SEIR_ABM, node_capacities, resolve_pars


class SimPack:
    """What a batch needs of one simulation: its tables (agent columns may live in a
    shared buffer instead), its lp_sim_params record and its initial results arrays."""

    def __init__(self, tables, sim_params, results, sim=None):
        self.tables, self.sim_params, self.results, self.sim = tables, sim_params, results, sim

    @classmethod
    def of(cls, sim):
        if not isinstance(sim, SEIR_ABM) or not sim._fused_ok():
            raise ValueError("every simulation must be a SEIR_ABM with the stock components")
        ds = sim._instance("DiseaseState_ABM")
        if ds is not None and len(ds.seed_schedule):
            raise ValueError("seed_schedule injections are not supported in a batch")
        if sim.t != 0:
            raise ValueError("simulations must not have been run")
        tb = sim.export_tables()
        return cls(tb, sim.sim_params(tb), sim._results_arrays(), sim)


class BatchedSims:
    def __init__(self, sims, tiles_per_block=0, timing=False, share_networks=False):
        if not sims:
            raise ValueError("BatchedSims needs at least one simulation")
        self._init([SimPack.of(s) for s in sims], None, tiles_per_block, timing, share_networks)

    @classmethod
    def from_packs(cls, packs, cols=None, tiles_per_block=0, timing=False, owner=None, share_networks=False):
        """Batch of SimPacks; `cols` = the already concatenated agent columns, if any."""
        self = cls.__new__(cls)
        self._init(list(packs), cols, tiles_per_block, timing, share_networks)
        self._owner = owner
        return self

    def _init(self, packs, cols, tiles_per_block, timing, share_networks=False):
        self.share_networks = bool(share_networks)
        self.packs = packs
        self.sims = [p.sim for p in packs]
        self.tables = [p.tables for p in packs]
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
        self._owner = None
        self.device_init = None    # (specs, infected slots per simulation) when drawn on the GPU
        self._net_inputs = None    # device populations and distances for set_gravity
        self._pack(cols)
        self.engine = None

    # -- packing --------------------------------------------------------------------
    def _pack(self, cols):
        tbs, S = self.tables, len(self.tables)
        nt, K = tbs[0].nt, tbs[0].n_strains
        caps = np.array([tb.capacity for tb in tbs], np.int64)
        nns = np.array([tb.n_nodes for tb in tbs], np.int64)
        self.slot_off = np.concatenate([[0], np.cumsum(caps)])
        self.node_off = np.concatenate([[0], np.cumsum(nns)])
        # network blocks: one per simulation, or with share_networks one per distinct network
        net_off = np.zeros(S, np.int64)
        blocks, first_of, pos = [], {}, 0
        for i, tb in enumerate(tbs):
            if self.share_networks:
                key = (int(nns[i]), hash(np.ascontiguousarray(tb.network_t).tobytes()))
                j = first_of.get(key)
                if j is not None and np.array_equal(tbs[j].network_t, tb.network_t):
                    net_off[i] = net_off[j]
                    continue
                first_of[key] = i
            net_off[i] = pos
            blocks.append(tb.network_t)
            pos += int(nns[i] * nns[i])
        self.net_off = net_off          # each simulation's first network_t entry
        N = int(self.node_off[-1])
        if cols is None:
            cols = {name: np.concatenate([tb.cols[name] for tb in tbs]) for name in AGENT_COLUMNS}
        arrays = dict(cols)
        arrays["seg_begin"] = np.concatenate([tb.seg_begin + o for tb, o in zip(tbs, self.slot_off)])
        arrays["live_end"] = np.concatenate([tb.live_end + o for tb, o in zip(tbs, self.slot_off)])
        arrays["seg_end"] = np.concatenate([tb.seg_end + o for tb, o in zip(tbs, self.slot_off)])
        arrays["node_sim"] = np.repeat(np.arange(S, dtype=np.int32), nns)
        for key, attr in (("r0_scalar", "r0s"), ("keep", "keep"), ("birth_rate_step", "birth_rate_step"),
                          ("thr_ri", "thr_ri"), ("thr_ipv", "thr_ipv")):
            arrays[key] = np.concatenate([getattr(tb, attr) for tb in tbs])
        sp = np.concatenate([p.sim_params for p in self.packs])
        sp["slot_base"] = self.slot_off[:-1]
        sp["net_off"] = net_off
        sp["node_begin"] = self.node_off[:-1]
        arrays["sim_params"] = sp
        arrays["network_t"] = np.concatenate(blocks)
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
        for bind, _ in RESULT_ARRAYS.values():
            arrays[bind] = np.ascontiguousarray(np.concatenate([p.results[bind] for p in self.packs], axis=1))
        self.arrays = arrays
        self.cfg = engmod.LPConfig(capacity=int(self.slot_off[-1]), n_nodes=N, n_sims=S, n_strains=K, nt=nt,
                                   n_sia_events=E, tiles_per_block=self.tiles_per_block,
                                   mix_mode=int(getattr(tbs[0], "mix_mode", 0) or 0))

    # -- running ----------------------------------------------------------------------
    def prepare(self):
        """Upload the packed buffers, or draw the agent columns on the GPU (not timed by run())."""
        if self.engine is None:
            if self.device_init is None:
                self.engine = engmod.Engine(self.cfg, self.arrays, timing=self.timing)
            else:
                specs, infected = self.device_init
                self.engine = engmod.Engine(self.cfg, self.arrays, timing=self.timing, device_only=set(AGENT_COLUMNS))
                slots = [np.asarray(x, np.int64) + self.slot_off[k] for k, x in enumerate(infected)]
                self.engine.init_population(specs, np.concatenate(slots) if slots else np.zeros(0, np.int64))
        return self.engine

    def set_gravity(self, k=None, a=None, b=None, c=None, max_rowsum=None, pull=True):
        """Rebuild every simulation's gravity network on the device (lp_build_network).
        Each of k (gravity_k * 10^gravity_k_exponent), a, b, c and max_rowsum (max_migr_frac)
        is None (the simulation's current value), a scalar, or one value per simulation.
        With `pull` the host copies of network_t and keep (and each simulation's tables)
        are refreshed, so a simulation still replays alone with its device-built network."""
        import torch

        S = len(self.tables)
        if len(set(self.net_off.tolist())) < S:
            raise ValueError("set_gravity builds one network per simulation; this batch shares networks "
                             "(share_networks=True)")
        if any(getattr(tb, "gravity_spec", None) is None for tb in self.tables):
            raise ValueError("set_gravity needs gravity networks (migration_method 'gravity') in every simulation")
        cols = []
        for i, v in enumerate((k, a, b, c, max_rowsum)):
            cur = [tb.gravity_spec[i] for tb in self.tables]
            cols.append(cur if v is None else np.broadcast_to(np.asarray(v, dtype=np.float64), (S,)).tolist())
        specs = [tuple(float(x) for x in row) for row in zip(*cols)]
        eng = self.prepare()
        if self._net_inputs is None:
            pop = np.concatenate([np.asarray(tb.init_pop, np.float64) for tb in self.tables])
            dist = np.concatenate([np.asarray(tb.distances, np.float64).reshape(-1) for tb in self.tables])
            self._net_inputs = (torch.as_tensor(pop, device="cuda"), torch.as_tensor(dist, device="cuda"))
        eng.build_network(specs, *self._net_inputs)
        for tb, spec in zip(self.tables, specs):
            tb.gravity_spec = spec
        if pull:
            eng.pull("network_t", self.arrays["network_t"])
            eng.pull("keep", self.arrays["keep"])
            for i, tb in enumerate(self.tables):
                n = tb.n_nodes
                tb.network_t = self.arrays["network_t"][self.net_off[i]:self.net_off[i] + n * n].copy()
                tb.keep = self.arrays["keep"][self.node_off[i]:self.node_off[i + 1]].copy()
        return specs

    def pull_sim(self, k, names=None):
        """Agent columns (sim-local slots) and results of simulation k from the device."""
        eng = self.prepare()
        eng.sync()      # the device-held records written back into the columns
        a, b = self.slot_off[k], self.slot_off[k + 1]
        na, nb = self.node_off[k], self.node_off[k + 1]
        cols, res = {}, {}
        for name in (names or AGENT_COLUMNS):
            cols[name] = eng.tensors[name][a:b].cpu().numpy().view(AGENT_COLUMNS[name][0])
        for attr, (bind, per) in RESULT_ARRAYS.items():
            full = eng.tensors[bind].cpu().numpy().reshape(self.arrays[bind].shape)
            res[attr] = full[:, na:nb]
        live = np.empty_like(self.arrays["live_end"])
        eng.pull("live_end", live)
        return cols, res, live[na:nb] - a

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
        if any(s is None for s in self.sims):
            raise ValueError("this batch was built without per-simulation objects")
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

    def subset(self, ks, tiles_per_block=None):
        """A new batch of simulations `ks` (their packed columns copied out of this one, or,
        for a batch drawn on the GPU, their initialisation specs)."""
        ks = list(ks)
        packs = [self.packs[k] for k in ks]
        tpw = self.tiles_per_block if tiles_per_block is None else tiles_per_block
        if self.device_init is not None:
            total = int(sum(self.slot_off[k + 1] - self.slot_off[k] for k in ks))
            cols = {name: np.empty(total, dtype=dt) for name, (dt, _) in AGENT_COLUMNS.items()}
            sub = BatchedSims.from_packs(packs, cols, tiles_per_block=tpw, timing=self.timing,
                                         share_networks=self.share_networks)
            sub.device_init = ([self.device_init[0][k] for k in ks], [self.device_init[1][k] for k in ks])
            return sub
        cols = {name: np.concatenate([self.arrays[name][self.slot_off[k]:self.slot_off[k + 1]] for k in ks])
                for name in AGENT_COLUMNS}
        return BatchedSims.from_packs(packs, cols, tiles_per_block=tpw, timing=self.timing,
                                      share_networks=self.share_networks)

    # 0: the library's default (the workgroups resident at once)
    GRID_CANDIDATES = {"ei_grid": (0, 1024, 2048, 4096), "sample_grid": (0, 1024, 2048, 4096)}
    TPW_DEFAULT = 16
    # a trial of a few early days of a few simulations cannot separate candidates closer
    # than this: the default stays unless another candidate beats it by more
    TUNE_MARGIN = 0.03

    def _time_subset(self, ks, days, use_graph, tpw, grids):
        import torch

        sub = self.subset(ks, tiles_per_block=tpw)
        for name, g in grids.items():
            setattr(sub.cfg, name, int(g))
        eng = sub.prepare()
        eng.run_days(0, 2, use_graph=use_graph)
        eng.sync()
        torch.cuda.synchronize()
        t = time.perf_counter()
        eng.run_days(2, min(self.cfg.nt, 2 + days), use_graph=use_graph)
        eng.sync()
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        sub.close()
        return el

    @classmethod
    def _pick(cls, times, default):
        """The fastest candidate if it beats `default` by more than TUNE_MARGIN, else `default`."""
        best = min(times, key=times.get)
        if default in times and times[best] >= times[default] * (1.0 - cls.TUNE_MARGIN):
            return default
        return best

    def autotune(self, candidates=(4, 8, 16), days=20, max_sims=32, use_graph=True, grids=False):
        """Pick tiles_per_block by timing `days` simulated days of the first `max_sims`
        simulations with each candidate (fresh contexts, so nothing of this batch runs);
        with `grids`, then the workgroup counts of the list pass and the round-1 sampler,
        one after the other from GRID_CANDIDATES (the launch shapes of lp_config).  A
        candidate replaces the default (16 tiles, the library's grids) only when it is
        more than TUNE_MARGIN faster.  Returns (tiles_per_block, times by tiles_per_block),
        and records the chosen grids in self.cfg and the trial times in self.tune_log."""
        ks = range(min(max_sims, len(self.packs)))
        times = {tpw: self._time_subset(ks, days, use_graph, tpw, {}) for tpw in candidates}
        best = self._pick(times, self.TPW_DEFAULT)
        self.tiles_per_block = best
        self.cfg.tiles_per_block = best
        self.tune_log = {"tiles_per_block": dict(times)}
        if grids:
            chosen = {}
            for name, cands in self.GRID_CANDIDATES.items():
                tried = {g: self._time_subset(ks, days, use_graph, best, {**chosen, name: g}) for g in cands}
                chosen[name] = self._pick(tried, 0)
                self.tune_log[name] = tried
            for name, g in chosen.items():
                setattr(self.cfg, name, int(g))
        return best, times

    def close(self):
        if self.engine is not None:
            self.engine.close()
            self.engine = None
        if self._owner is not None:
            self._owner.release()
            self._owner = None


# -- parallel host initialisation of a large batch ---------------------------------------
class _SharedColumns:
    """One shared-memory block per agent column, sized for the whole batch."""

    def __init__(self, total):
        self.blocks, self.arrays = {}, {}
        try:
            for name, (dt, _) in AGENT_COLUMNS.items():
                nbytes = max(1, int(total) * np.dtype(dt).itemsize)
                blk = shared_memory.SharedMemory(create=True, size=nbytes)
                self.blocks[name] = blk
                self.arrays[name] = np.ndarray((int(total),), dtype=dt, buffer=blk.buf)
        except BaseException:
            self.release()
            raise

    def release(self):
        self.arrays = {}
        for blk in self.blocks.values():
            try:
                blk.close()
                blk.unlink()
            except FileNotFoundError:
                pass
        self.blocks = {}


_W = {}


def _worker_init(names, total, slot_off, build):
    _W["blocks"] = {n: shared_memory.SharedMemory(name=b) for n, b in names.items()}
    _W["arrays"] = {n: np.ndarray((total,), dtype=AGENT_COLUMNS[n][0], buffer=_W["blocks"][n].buf) for n in names}
    _W["slot_off"], _W["build"] = slot_off, build


def _worker_build(args):
    k, pars = args
    sim = _W["build"](pars)
    tb = sim.export_tables()
    a, b = _W["slot_off"][k], _W["slot_off"][k + 1]
    if tb.capacity != b - a:
        raise RuntimeError(f"simulation {k}: capacity {tb.capacity} != planned {b - a}")
    for name in AGENT_COLUMNS:
        _W["arrays"][name][a:b] = tb.cols[name]
    sp, res = sim.sim_params(tb), sim._results_arrays()
    tb.cols = None
    return k, tb, sp, res


def build_replicates(pars_list, build, workers=8, tiles_per_block=0, timing=False, log=None):
    """BatchedSims of `build(pars)` for every pars in `pars_list`, initialised in `workers`
    forked processes that write the agent columns into shared memory.  Must run before the
    calling process touches the GPU."""
    caps = []
    for p in pars_list:
        rp = resolve_pars(p)
        pop = np.atleast_1d(rp.init_pop).astype(np.int64)
        caps.append(int(PeopleFrame(pop, node_capacities(rp)).capacity))
    slot_off = np.concatenate([[0], np.cumsum(caps)]).astype(np.int64)
    total = int(slot_off[-1])
    shared = _SharedColumns(total)
    try:
        names = {n: blk.name for n, blk in shared.blocks.items()}
        packs = [None] * len(pars_list)
        ctx = mp.get_context("fork")
        with ctx.Pool(workers, initializer=_worker_init, initargs=(names, total, slot_off, build)) as pool:
            for done, (k, tb, sp, res) in enumerate(pool.imap_unordered(_worker_build, enumerate(pars_list)), 1):
                packs[k] = SimPack(tb, sp, res)
                if log is not None and (done % 32 == 0 or done == len(pars_list)):
                    log(f"[batch] initialised {done}/{len(pars_list)} simulations")
        return BatchedSims.from_packs(packs, shared.arrays, tiles_per_block=tiles_per_block, timing=timing, owner=shared)
    except BaseException:
        shared.release()
        raise


def build_replicates_device(pars_list, build, tiles_per_block=0, timing=False, log=None, share_networks=False):
    """BatchedSims of `build(pars)` for every pars in `pars_list`, their agent columns drawn
    on the GPU when the batch is prepared (the host builds only node tables, schedules and
    the exact-count initial infections)."""
    packs, specs, infected = [], [], []
    for k, p in enumerate(pars_list):
        p.init_on_device = True
        sim = build(p)
        tb = sim._export_tables()
        tb.cols = None
        packs.append(SimPack(tb, sim.sim_params(tb), sim._results_arrays()))
        specs.append(sim.device_init_spec())
        infected.append(np.asarray(sim.init_spec.get("infected", np.zeros(0, np.int64)), np.int64))
        if log is not None and ((k + 1) % 128 == 0 or k + 1 == len(pars_list)):
            log(f"[batch] built {k + 1}/{len(pars_list)} simulations (population drawn on the GPU)")
    total = int(sum(pk.tables.capacity for pk in packs))
    cols = {name: np.empty(total, dtype=dt) for name, (dt, _) in AGENT_COLUMNS.items()}
    batch = BatchedSims.from_packs(packs, cols, tiles_per_block=tiles_per_block, timing=timing,
                                   share_networks=share_networks)
    batch.device_init = (specs, infected)
    return batch
