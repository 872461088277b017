"""SEIR_ABM and its components -- the host side of the MI355X daily step.

Drop-in for the component API of src/laser_polio/model.py:
  * ``SEIR_ABM(pars)``, ``sim.components = [...]`` (instantiated and ordered by
    class name against ``default_run_order``, model.py:209-244), ``sim.run()``,
    ``sim.people``, ``sim.results``, ``sim.t/nt/datevec/nodes/should_stop``;
  * components ``VitalDynamics_ABM``, ``DiseaseState_ABM``, ``RI_ABM``,
    ``SIA_ABM``, ``Transmission_ABM`` with ``__init__(sim)``, ``step()``,
    ``log(t)``, ``plot()`` and the same result arrays (names, shapes, int32).

Initialisation (ages, death dates, timers, immunity, seeding, heterogeneity,
network) runs on the host in numpy, as in the reference.  Every simulated day
runs on the GPU through liblpstep.so: ``run()`` launches all days fused (one
hipGraph per day); calling a component's ``step()`` directly runs that stage
alone.  Agents are kept node-segmented (see laserframe.PeopleFrame): node n's
agents occupy slots [seg_begin[n], live_end[n]), initial agents first, newborns
appended in birth order, which is the within-node order the reference's
infection sampling depends on (model.py:1055-1061).
"""

import numbers
from collections import defaultdict
from copy import deepcopy
from datetime import datetime

import numpy as np

from . import distributions as lpd
from . import engine as engmod
from .laserframe import LaserFrame, PeopleFrame
from .network import distance, gravity, radiation, row_normalizer
from .pars import default_pars, default_run_order
from .propertyset import PropertySet
from .utils import TimingStats, calc_capacity, create_cumulative_deaths, daterange, seasonality_of
from .utils import date as to_date

__all__ = ["RI_ABM", "SEIR_ABM", "SIA_ABM", "DiseaseState_ABM", "Transmission_ABM", "VitalDynamics_ABM", "SimTables"]

INT32_MAX = np.iinfo(np.int32).max

# name -> (dtype, default) of every agent column the step reads or writes.
AGENT_COLUMNS = {
    "disease_state": (np.int8, -1),
    "strain": (np.int8, 0),
    "node_id": (np.int16, -1),
    "exposure_timer": (np.uint16, 0),
    "infection_timer": (np.uint16, 0),
    "paralysis_timer": (np.uint16, 0),
    "potentially_paralyzed": (np.int8, -1),
    "paralyzed": (np.int8, 0),
    "ipv_protected": (np.int8, 0),
    "chronically_missed": (np.uint8, 0),
    "acq_risk_multiplier": (np.float32, 1.0),
    "daily_infectivity": (np.float32, 1.0),
    "date_of_birth": (np.int32, -1),
    "date_of_death": (np.int32, INT32_MAX),
    "ri_timer": (np.int16, -1),
}

# results attribute -> (binding name, per-strain)
RESULT_ARRAYS = {
    "S": ("S", False),
    "E": ("E", False),
    "I": ("I", False),
    "R": ("R", False),
    "E_by_strain": ("E_by_strain", True),
    "I_by_strain": ("I_by_strain", True),
    "potentially_paralyzed": ("potentially_paralyzed_count", False),
    "paralyzed": ("paralyzed_count", False),
    "new_potentially_paralyzed": ("new_potentially_paralyzed", False),
    "new_paralyzed": ("new_paralyzed", False),
    "pop": ("pop", False),
    "births": ("births", False),
    "deaths": ("deaths", False),
    "new_exposed": ("new_exposed", False),
    "new_exposed_by_strain": ("new_exposed_by_strain", True),
    "ri_vaccinated": ("ri_vaccinated", False),
    "ri_protected": ("ri_protected", False),
    "ri_new_exposed_by_strain": ("ri_new_exposed_by_strain", True),
    "ipv_vaccinated": ("ipv_vaccinated", False),
    "sia_vaccinated": ("sia_vaccinated", False),
    "sia_protected": ("sia_protected", False),
    "sia_new_exposed_by_strain": ("sia_new_exposed_by_strain", True),
}

STAGE_OF = {
    "VitalDynamics_ABM": engmod.STAGE_VITAL,
    "DiseaseState_ABM": engmod.STAGE_DISEASE,
    "RI_ABM": engmod.STAGE_RI,
    "SIA_ABM": engmod.STAGE_SIA,
    "Transmission_ABM": engmod.STAGE_TRANSMISSION,
}


def _vaccine_strain(vtype):
    """Strain given by a vaccine type: nOPV -> 2, mOPV/tOPV -> 1, else 1 (model.py:1935-1940, 2115-2120)."""
    if "nOPV" in vtype:
        return 2
    return 1


def dist_spec(dist):
    """lp_dist of a duration Distribution (the device samplers of lp_init_population)."""
    d, p = dist.dist_type, dist.pars
    kinds = {"constant": (engmod.LP_DIST_CONSTANT, p.get("value", 1), 0.0),
             "normal": (engmod.LP_DIST_NORMAL, p.get("mean", 0.0), p.get("std", 1.0)),
             "gamma": (engmod.LP_DIST_GAMMA, p.get("shape", 2.0), p.get("scale", 1.0)),
             "lognormal": (engmod.LP_DIST_LOGNORMAL, p.get("mean", 1.0), p.get("sigma", 0.5)),
             "poisson": (engmod.LP_DIST_POISSON, p.get("lam", 5), 0.0),
             "exponential": (engmod.LP_DIST_EXPONENTIAL, p.get("scale", 1.0), 0.0),
             "uniform": (engmod.LP_DIST_UNIFORM_INT, p.get("min", 2), p.get("max", 10))}
    if d not in kinds:
        raise ValueError(f"device initialisation has no sampler for the {d} distribution")
    k, a, b = kinds[d]
    return engmod.LPDist(k, float(a), float(b))


def timer_from_samples(x, hi=127):
    """Float duration samples -> timers: truncate toward zero, clip to [0, hi] (model.py:575-579)."""
    return np.clip(np.trunc(np.asarray(x, dtype=np.float64)), 0, hi).astype(np.uint16)


def load_sim(path, pars=None, verbose=1):
    """A model continued from a snapshot, as run_sim.py:388-405 builds it: the agents of
    `path` (PeopleFrame.load_snapshot, capacities for pars' init_pop and cbr over dur + 10
    days), every stock component made by its init_from_file, results.R from the snapshot,
    pars.old_r0 from the snapshot's r0 (daily_infectivity rescaled when pars.r0 differs).
    The components run in the stock order (default_run_order)."""
    rp = resolve_pars(pars)
    people, results_r, loaded = PeopleFrame.load_snapshot(path, n_ppl=rp.init_pop, cbr=rp.cbr, nt=rp.dur + 10)
    sim = SEIR_ABM.init_from_file(people, pars, verbose)
    if loaded and loaded.get("r0") is not None:
        sim.pars.old_r0 = loaded["r0"]
    made = {}
    for cls in (DiseaseState_ABM, VitalDynamics_ABM, SIA_ABM, RI_ABM, Transmission_ABM):
        made[cls.__name__] = cls.init_from_file(sim)
    if results_r is not None:
        R = sim.results.raw("R")
        k = min(len(R), len(results_r))
        R[:k] = results_r[:k]
    order = [made[name] for name in default_run_order if name in made]
    sim._components = [type(c) for c in order]
    sim.instances = order
    return sim


class SimTables:
    """Everything one simulation hands to the GPU step (and to the oracle):
    agent columns, node tables, sim scalars, network, mortality table, schedule."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


# =============================================================================
def node_capacities(pars, frame_pop=None):
    """Agent slots per node for a run: the agents of the frame (`frame_pop`, default the
    initial population) plus the births that calc_capacity expects of the node's whole
    population over dur + 100 days, with a 4-sigma margin (newborns of a node stay in its
    segment)."""
    pop = np.atleast_1d(pars.init_pop).astype(np.int64)
    fp = pop if frame_pop is None else np.atleast_1d(frame_pop).astype(np.int64)
    cbr = None if pars.cbr is None else np.broadcast_to(np.atleast_1d(np.asarray(pars.cbr, dtype=np.float64)), pop.shape)
    cap = np.array(fp, dtype=np.int64)
    if cbr is not None:
        for n in range(len(pop)):
            eb = calc_capacity(pop[n], pars.dur + 100, cbr[n]) - pop[n]
            if eb > 0:
                cap[n] = int((1 + 4 / np.sqrt(eb)) * (fp[n] + eb)) + 16
    return cap


def resolve_pars(pars):
    """The reference defaults overridden by `pars` (unknown keys dropped), as SEIR_ABM uses them."""
    out = deepcopy(default_pars)
    if pars is not None:
        if not isinstance(pars, PropertySet):
            pars = PropertySet(pars)
        out <<= {k: v for k, v in pars.items() if k in out}
    return out


class SEIR_ABM:
    """Agent-based SEIR model of polio; disease_state codes -1 dead/unborn, 0 S, 1 E, 2 I, 3 R."""

    def common_init(self, pars, verbose):
        self.perf_stats = TimingStats()
        self.pars = deepcopy(default_pars)
        if pars is not None:
            if not isinstance(pars, PropertySet):
                pars = PropertySet(pars)
            unexpected = set(pars.keys()) - set(self.pars.keys())
            filtered = {k: v for k, v in pars.items() if k in self.pars}
            if unexpected and (pars.get("verbose", 1) or 0) >= 1:
                print(f"Warning: ignoring unexpected parameters: {sorted(unexpected)}")
            self.pars <<= filtered
        pars = self.pars
        self.verbose = pars["verbose"] if "verbose" in pars else verbose
        if pars.seed is None:
            now = datetime.now()  # noqa: DTZ005
            pars.seed = now.microsecond ^ int(now.timestamp())
        self.rng = lpd.set_seed(pars.seed)
        self.t = 0
        self.nt = pars.dur + 1
        self.datevec = daterange(pars.start_date, days=self.nt)
        self.should_stop = False

    def __init__(self, pars: PropertySet = None, verbose=1):
        self.perf_stats = TimingStats()
        with self.perf_stats.start("SEIR_ABM.__init__()"):
            self.common_init(pars, verbose)
            pars = self.pars
            pars.init_pop = np.atleast_1d(pars.init_pop).astype(np.int64)
            if pars.init_sus_by_age is not None:
                # only the susceptibles are agents; the immune rest of each node is counted
                # in results.R (model.py:115-121, 614-665)
                if pars.init_on_device:
                    raise ValueError("init_sus_by_age is initialised on the host (init_on_device must be False)")
                by_node = pars.init_sus_by_age.groupby("node_id")["n_susceptible"].sum()
                sus = np.zeros(len(pars.init_pop), dtype=np.int64)
                sus[by_node.index.to_numpy(dtype=np.int64)] = by_node.to_numpy().astype(np.int64)
                pars.init_sus = sus
                pop = sus
            else:
                pop = pars.init_pop
            if np.any(pop <= 0) or np.any(np.isnan(pop.astype(np.float64))):
                raise ValueError("pop_by_node must be positive & non-nan")
            self.device_init = bool(pars.init_on_device)
            self.people = PeopleFrame(pop, node_capacities(pars, frame_pop=pop), lazy=self.device_init)
            p = self.people
            p.add_scalar_property("disease_state", dtype=np.int8, default=-1)
            if not self.device_init:
                for n in range(len(pop)):
                    p.raw("disease_state")[p.initial_slice(n)] = 0
            p.add_scalar_property("potentially_paralyzed", dtype=np.int8, default=-1)
            p.add_scalar_property("paralyzed", dtype=np.int8, default=0)
            p.add_scalar_property("ipv_protected", dtype=np.int8, default=0)
            p.add_scalar_property("strain", dtype=np.int8, default=0)
            p.add_scalar_property("chronically_missed", dtype=np.uint8, default=0)
            n0 = int(pop.sum())
            n_missed = int(pars.missed_frac * n0)
            if n_missed > 0 and not self.device_init:
                p.raw("chronically_missed")[p.initial_to_slot(self.rng.choice(n0, size=n_missed, replace=False))] = 1
            p.add_scalar_property("node_id", dtype=np.int16, default=-1)
            if not self.device_init:
                for n in range(len(pop)):
                    p.raw("node_id")[p.initial_slice(n)] = n
            self.results = LaserFrame(capacity=1)
            self.nodes = np.arange(len(pop))
            self._init_bookkeeping()

    def _init_bookkeeping(self):
        self._components = []
        self.instances = []
        self._engine = None
        self._touched = set()
        self._dev_newer = set()
        self._timing = False
        self._device_initialized = False
        self.init_spec = {"missed_frac": float(self.pars.missed_frac)}   # device initialisation inputs

    @classmethod
    def init_from_file(cls, people, pars=None, verbose=1):
        """A model around agents loaded from a snapshot (PeopleFrame.load_snapshot;
        model.py:177-195): nothing is drawn for the loaded agents.  Components are then
        made with their own init_from_file (see load_sim)."""
        model = cls.__new__(cls)
        model.common_init(pars, verbose)
        pars = model.pars
        if pars.init_pop is None:
            pars.init_pop = np.asarray(people.init_pop, dtype=np.int64)
        pars.init_pop = np.atleast_1d(pars.init_pop).astype(np.int64)
        if len(pars.init_pop) != people.n_nodes:
            raise ValueError(f"init_pop has {len(pars.init_pop)} nodes, the snapshot {people.n_nodes}")
        model.device_init = False
        model.people = people
        model.results = LaserFrame(capacity=1)
        model.nodes = np.arange(people.n_nodes)
        model._init_bookkeeping()
        return model

    # -- components -------------------------------------------------------------
    @property
    def components(self) -> list:
        return self._components

    @components.setter
    def components(self, components: list) -> None:
        lookup = {c.__name__: c for c in components}
        ordered = [lookup[name] for name in default_run_order if name in lookup]
        self._components = ordered
        self.instances = []
        for cls in ordered:
            with self.perf_stats.start(cls.__name__ + ".__init__()"):
                self.instances.append(cls(self))
        self._drop_engine()
        if self.device_init and not self._device_initialized:
            # the population is drawn on the GPU the first time an agent column is read
            object.__setattr__(self.people, "_sync_hook", self._draw_hook)

    def _draw_hook(self, frame, name):
        if self._binding_of(frame, name) is not None and self._engine is None:
            self._ensure_engine()
        self._sync_hook(frame, name)

    def _instance(self, name):
        for c in self.instances:
            if type(c).__name__ == name:
                return c
        return None

    # -- host <-> device ----------------------------------------------------------
    def _binding_of(self, frame, name):
        if frame is self.people:
            return name if name in AGENT_COLUMNS else None
        if frame is self.results and name in RESULT_ARRAYS:
            return RESULT_ARRAYS[name][0]
        return None

    def _sync_hook(self, frame, name):
        bind = self._binding_of(frame, name)
        if bind is None or self._engine is None:
            return
        key = (id(frame), name)
        if key in self._dev_newer:
            self._engine.pull(bind, frame.raw(name))
            self._dev_newer.discard(key)
        self._touched.add(key)

    def _push_touched(self):
        for frame in (self.people, self.results):
            for name in frame.column_names():
                key = (id(frame), name)
                bind = self._binding_of(frame, name)
                if bind is not None and key in self._touched and key not in self._dev_newer:
                    self._engine.push(bind, frame.raw(name))
        self._touched.clear()

    def _mark_device_newer(self):
        for frame in (self.people, self.results):
            for name in frame.column_names():
                if self._binding_of(frame, name) is not None:
                    self._dev_newer.add((id(frame), name))
        self._engine.pull("live_end", self.people.live_end)

    def _drop_engine(self):
        eng = getattr(self, "_engine", None)
        if eng is not None:
            self._pull_all()
            eng.close()
        self._engine = None
        for frame in (getattr(self, "people", None), getattr(self, "results", None)):
            if frame is not None:
                object.__setattr__(frame, "_sync_hook", None)

    def _pull_all(self):
        for frame in (self.people, self.results):
            for name in list(frame.column_names()):
                key = (id(frame), name)
                if key in self._dev_newer:
                    self._engine.pull(self._binding_of(frame, name), frame.raw(name))
                    self._dev_newer.discard(key)

    # -- exact checkpoints --------------------------------------------------------------
    def save_checkpoint(self, path):
        """Everything a run continues from, to `path` (.npz): every agent column over the
        whole capacity (the values drawn for slots not yet handed out included), the node
        segments, every results array, the next tick, the host generator's state (seed
        schedule draws) and the parameters.  Draws are keyed by (seed, slot or node, day),
        so a model built with the same parameters and components that loads it
        (load_checkpoint) and runs on ends with the same state as a run never stopped."""
        import json

        from .laserframe import _jsonable

        if self._engine is not None:
            self._engine.sync()
            self._pull_all()
        p, r = self.people, self.results
        arrays = {"seg_begin": p.seg_begin, "seg_end": p.seg_end, "live_end": p.live_end}
        for name in p.column_names():
            arrays["col_" + name] = p.raw(name)
        for name in r.column_names():
            arrays["res_" + name] = r.raw(name)
        meta = {"format": "laser_polio_amd.checkpoint/1", "t": int(self.t), "nt": int(self.nt),
                "capacity": int(p.capacity), "n_nodes": int(p.n_nodes), "rng": self.rng.bit_generator.state,
                "pars": _jsonable(self.pars)}
        arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        with open(path, "wb") as fh:
            np.savez(fh, **arrays)

    def load_checkpoint(self, path):
        """Continue from a checkpoint saved by a model with the same parameters and
        components (same capacity, nodes and duration): its agent columns, node segments,
        results, tick and host generator state replace this model's; the next run()
        recounts every carried statistic from the loaded columns and goes on from the saved
        tick."""
        import json

        with np.load(path, allow_pickle=False) as z:
            meta = json.loads(bytes(z["meta"]).decode())
            p, r = self.people, self.results
            if (meta["capacity"], meta["n_nodes"], meta["nt"]) != (p.capacity, p.n_nodes, self.nt):
                raise ValueError("checkpoint of a model with another capacity, node count or duration")
            if not (np.array_equal(z["seg_begin"], p.seg_begin) and np.array_equal(z["seg_end"], p.seg_end)):
                raise ValueError("checkpoint of a model with other node segments")
            eng = self._engine
            if eng is not None:              # nothing of the device state survives the load
                eng.close()
                self._engine = None
                for frame in (self.people, self.results):
                    object.__setattr__(frame, "_sync_hook", None)
            for name in p.column_names():
                p.raw(name)[...] = z["col_" + name]
            for name in r.column_names():
                r.raw(name)[...] = z["res_" + name]
            p.live_end[...] = z["live_end"]
        self.t = int(meta["t"])
        self.rng.bit_generator.state = meta["rng"]
        self._device_initialized = True     # the loaded columns are the population
        self.should_stop = False

    def export_tables(self):
        """SimTables of the current host state (used by the GPU engine and by the oracle).
        With init_on_device the population is drawn on the GPU first and pulled."""
        if self.device_init:
            self._ensure_engine()
            self._pull_all()
        return self._export_tables()

    def _export_tables(self):
        pars, p = self.pars, self.people
        N, cap, K = len(self.nodes), p.capacity, len(pars.strain_ids)
        vd, ds, ri, sia, tx = (self._instance(n) for n in default_run_order)
        cols = {}
        for name, (dt, default) in AGENT_COLUMNS.items():
            cols[name] = p.raw(name) if p.has(name) else np.full(cap, default, dtype=dt)
        r0s = np.broadcast_to(np.asarray(tx.r0_scalars if tx else 1.0, dtype=np.float64), (N,)).astype(np.float32)
        if tx is not None:
            net = np.asarray(tx.network, dtype=np.float64)
            keep = (1.0 - net.sum(axis=1)).astype(np.float32)
            network_t = np.ascontiguousarray(net.T.astype(np.float32)).reshape(-1)
            dist, gspec = getattr(tx, "distances", None), getattr(tx, "gravity_spec", None)
        else:
            keep = np.ones(N, np.float32)
            network_t = np.zeros(N * N, np.float32)
            dist, gspec = None, None
        if vd is not None:
            # the reference's step_size * birth_rate (model.py:1712), kept in float64
            brs = vd.step_size * np.asarray(vd.birth_rate, dtype=np.float64)
            km_cd = vd.km_cd
        else:
            brs = np.zeros(N, np.float64)
            km_cd = create_cumulative_deaths(int(pars.init_pop.sum()), 100)
        ri_on = ri is not None and pars.vx_prob_ri is not None
        thr_ri = engmod.thr_from_prob64(np.broadcast_to(np.asarray(pars.vx_prob_ri if ri_on else 0.0, dtype=np.float64), (N,)))
        thr_ipv = engmod.thr_from_prob64(
            np.broadcast_to(np.asarray(pars.vx_prob_ipv if (ri_on and pars.vx_prob_ipv is not None) else 0.0, dtype=np.float64), (N,))
        )
        strain_r0 = np.zeros(engmod.MAX_STRAINS, np.float32)
        strain_r0[:K] = np.array(list(pars.strain_r0_scalars.values()), dtype=np.float64)[:K].astype(np.float32)
        zi = float(pars.node_seeding_zero_inflation)
        seed_days = sorted(ds.seed_schedule.keys()) if ds is not None else []
        sched = self._schedule(vd, ds, ri, sia, tx)
        return SimTables(
            cols=cols,
            capacity=cap,
            n_nodes=N,
            n_strains=K,
            nt=self.nt,
            seg_begin=p.seg_begin,
            live_end=p.live_end,
            seg_end=p.seg_end,
            r0s=r0s,
            keep=keep,
            network_t=network_t,
            distances=dist,
            gravity_spec=gspec,
            init_pop=np.asarray(pars.init_pop, dtype=np.float64),
            birth_rate_step=brs,
            thr_ri=thr_ri,
            thr_ipv=thr_ipv,
            km_cd=np.asarray(km_cd, dtype=np.int64),
            seed=int(pars.seed) & 0xFFFFFFFFFFFFFFFF,
            strain_r0=strain_r0,
            thr_paralysis=int(engmod.thr_from_prob64(np.float64(np.float32(pars.p_paralysis)))),
            thr_zi=int(engmod.thr_from_prob64(zi)),
            zi=np.float32(zi),
            disp_r=max(1, int(np.round(pars.node_seeding_dispersion))),
            ri_strain=_vaccine_strain(getattr(pars, "ri_vaccine_type", "tOPV") if "ri_vaccine_type" in pars else "tOPV"),
            ri_step=int(pars.step_size_RI_ABM),
            newborn_ri_timer=int(pars.ri_newborn_timer) if ri is not None else -1,
            last_seed_day=int(max(seed_days)) if seed_days else -1,
            mix_mode=int(pars.mix_mode or 0),
            **sched,
        )

    def _schedule(self, vd, ds, ri, sia, tx):
        pars, nt, N = self.pars, self.nt, len(self.nodes)
        flags = np.zeros(nt, np.int32)
        sia_begin = np.zeros(nt, np.int32)
        sia_count = np.zeros(nt, np.int32)
        season = np.zeros(nt, np.float32)
        ev_strain, ev_min, ev_max, ev_thr = [], [], [], []
        if tx is not None:
            flags[0] |= engmod.DAY_TX
        ri_on = ri is not None and pars.vx_prob_ri is not None
        events_by_day = defaultdict(list)
        if sia is not None and pars.vx_prob_sia is not None:
            day_index = {d: i for i, d in enumerate(self.datevec)}
            for ev in sia.sia_schedule:
                t = day_index.get(to_date(ev["date"]))
                if t is not None and t >= 1:
                    events_by_day[t].append(ev)
        for t in range(1, nt):
            f = 0
            if vd is not None:
                f |= engmod.DAY_VD_STEP if t % vd.step_size == 0 else engmod.DAY_VD_COPY
            if ds is not None:
                f |= engmod.DAY_DS
                if pars.stop_if_no_cases:
                    f |= engmod.DAY_STOPCHK
            if ri_on and t % ri.step_size == 0:
                f |= engmod.DAY_RI
            if events_by_day.get(t):
                f |= engmod.DAY_SIA
                sia_begin[t] = len(ev_strain)
                sia_count[t] = len(events_by_day[t])
                for ev in events_by_day[t]:
                    vtype = ev["vaccinetype"]
                    eff = pars["vx_efficacy"][vtype]
                    p32 = np.broadcast_to(np.asarray(pars["vx_prob_sia"], dtype=np.float32), (N,)).astype(np.float64)
                    mask = np.zeros(N, bool)
                    mask[np.asarray(ev["nodes"], dtype=np.int64)] = True
                    thr = np.zeros((N, 2), np.uint64)
                    thr[:, 0] = np.where(mask, engmod.thr_from_prob64(p32), 0)
                    thr[:, 1] = np.where(mask, engmod.thr_from_prob64(p32 * float(eff)), 0)
                    lo, hi = ev["age_range"]
                    ev_strain.append(_vaccine_strain(vtype))
                    ev_min.append(int(lo))
                    ev_max.append(int(hi))
                    ev_thr.append(thr)
            if tx is not None:
                f |= engmod.DAY_TX
                season[t] = np.float32(seasonality_of(self.datevec[t], pars["seasonal_amplitude"], pars["seasonal_peak_doy"]))
            flags[t] = f
        E = len(ev_strain)
        return {
            "day_flags": flags,
            "day_sia_begin": sia_begin,
            "day_sia_count": sia_count,
            "season": season,
            "sia_strain": np.array(ev_strain, np.int32),
            "sia_min_age": np.array(ev_min, np.int32),
            "sia_max_age": np.array(ev_max, np.int32),
            "sia_thr": (np.stack(ev_thr).reshape(-1) if E else np.zeros(0, np.uint64)),
            "n_sia_events": E,
        }

    def _results_arrays(self):
        out = {}
        N, K = len(self.nodes), len(self.pars.strain_ids)
        for attr, (bind, per_strain) in RESULT_ARRAYS.items():
            shape = (self.nt, N, K) if per_strain else (self.nt, N)
            if self.results.has(attr):
                out[bind] = self.results.raw(attr)
            else:
                a = np.zeros(shape, np.int32)
                if attr == "pop":
                    a[0] = self.pars.init_pop
                out[bind] = a
        return out

    def sim_params(self, tb):
        """lp_sim_params record of this simulation (slot / node / network offsets 0)."""
        sp = np.zeros(1, dtype=engmod.SIM_PARAMS_DTYPE)
        sp["seed"] = tb.seed
        sp["slot_base"] = 0
        sp["net_off"] = 0
        sp["node_begin"] = 0
        sp["n_nodes"] = tb.n_nodes
        sp["strain_r0"] = tb.strain_r0
        sp["thr_paralysis"] = tb.thr_paralysis
        sp["thr_zi"] = tb.thr_zi
        sp["zi"] = tb.zi
        sp["disp_r"] = tb.disp_r
        sp["ri_strain"] = tb.ri_strain
        sp["ri_step"] = tb.ri_step
        sp["newborn_ri_timer"] = tb.newborn_ri_timer
        sp["last_seed_day"] = tb.last_seed_day
        return sp

    def _ensure_engine(self):
        if self._engine is not None:
            self._push_touched()
            return self._engine
        tb = self._export_tables()
        sp = self.sim_params(tb)
        arrays = dict(tb.cols)
        arrays.update(
            seg_begin=tb.seg_begin,
            live_end=tb.live_end,
            seg_end=tb.seg_end,
            node_sim=np.zeros(tb.n_nodes, np.int32),
            r0_scalar=tb.r0s,
            keep=tb.keep,
            birth_rate_step=tb.birth_rate_step,
            thr_ri=tb.thr_ri,
            thr_ipv=tb.thr_ipv,
            sim_params=sp,
            network_t=tb.network_t,
            km_cd=tb.km_cd,
            day_flags=tb.day_flags,
            day_sia_begin=tb.day_sia_begin,
            day_sia_count=tb.day_sia_count,
            season=tb.season,
        )
        if tb.n_sia_events:
            arrays.update(sia_strain=tb.sia_strain, sia_min_age=tb.sia_min_age, sia_max_age=tb.sia_max_age, sia_thr=tb.sia_thr)
        arrays.update(self._results_arrays())
        cfg = engmod.LPConfig(
            capacity=tb.capacity,
            n_nodes=tb.n_nodes,
            n_sims=1,
            n_strains=tb.n_strains,
            nt=tb.nt,
            n_sia_events=tb.n_sia_events,
            tiles_per_block=int(self.pars.tiles_per_block or 0),
            mix_mode=int(self.pars.mix_mode or 0),
            ei_grid=int(self.pars.get("ei_grid") or 0),
            sample_grid=int(self.pars.get("sample_grid") or 0),
        )
        device_only = set(AGENT_COLUMNS) if (self.device_init and not self._device_initialized) else set()
        self._engine = engmod.Engine(cfg, arrays, timing=self._timing, device_only=device_only)
        for frame in (self.people, self.results):
            object.__setattr__(frame, "_sync_hook", self._sync_hook)
        self._touched.clear()
        self._dev_newer.clear()
        if device_only:
            self._engine.init_population([self.device_init_spec()], self.init_spec.get("infected", np.zeros(0, np.int64)))
            self._device_initialized = True
            for name in AGENT_COLUMNS:
                if self.people.has(name):
                    self._dev_newer.add((id(self.people), name))
        return self._engine

    def device_init_spec(self):
        """LPInitSim of this simulation from the components' initialisation inputs."""
        pars, spec = self.pars, self.init_spec
        s = engmod.LPInitSim()
        s.seed = int(np.random.default_rng([int(pars.seed) & 0xFFFFFFFF, 0x1D]).integers(1, 2**63))
        ages = spec.get("ages", ("exponential", 25 * 365.0, 80 * 365.0))
        if ages[0] == "exponential":
            s.age_model, s.age_mean_days, s.age_cap_days = 0, float(ages[1]), float(ages[2])
        else:
            cdf, lo, hi = ages[1], ages[2], ages[3]
            if len(cdf) > engmod.INIT_PYR:
                raise ValueError(f"device initialisation holds at most {engmod.INIT_PYR} age pyramid bins")
            s.age_model, s.n_pyr = 1, len(cdf)
            for b in range(len(cdf)):
                s.pyr_cdf[b], s.pyr_lo[b], s.pyr_hi[b] = float(cdf[b]), int(lo[b]), int(hi[b])
        s.init_immun = float(spec.get("init_immun", 0.0))
        bands = spec.get("bands", [])
        if len(bands) > engmod.INIT_BANDS:
            raise ValueError(f"device initialisation holds at most {engmod.INIT_BANDS} recovered-by-age bands")
        s.n_bands = len(bands)
        for b, (lo, hi, frac) in enumerate(bands):
            s.band_lo[b], s.band_hi[b], s.band_frac[b] = int(lo), int(hi), float(frac)
        for field in ("dur_exp", "dur_inf", "t_to_paralysis"):
            setattr(s, field, dist_spec(getattr(pars, field)))
        het = spec.get("het", (0, 0.0, 0.0, 1.0, 0.0))
        s.het_model, s.risk_mu, s.risk_sigma, s.inf_mean, s.rho = het[0], float(het[1]), float(het[2]), float(het[3]), float(het[4])
        ri = spec.get("ri")
        if ri is not None:
            s.ri, s.ri_lo, s.ri_hi = 1, float(ri[0]), float(ri[1])
        s.vd = 1 if spec.get("vd") else 0
        s.missed_frac = float(spec.get("missed_frac", 0.0))
        return s

    def _engine_step(self, mask):
        eng = self._ensure_engine()
        eng.step(self.t, mask)
        self._mark_device_newer()

    def _fused_ok(self):
        stock = {VitalDynamics_ABM, DiseaseState_ABM, RI_ABM, SIA_ABM, Transmission_ABM}
        return all(type(c) in stock for c in self.instances)

    # -- running ------------------------------------------------------------------
    def run(self, use_graph=True, until=None):
        """Run ticks self.t .. nt - 1 (with `until`, stop before tick `until`: a later run()
        goes on from there, e.g. after save_checkpoint)."""
        if self.verbose >= 1:
            print("Initialization complete. Running simulation...")
        end = self.nt if until is None else max(self.t, min(int(until), self.nt))
        with self.perf_stats.start("SEIR_ABM.run()"):
            if self._fused_ok():
                self._run_fused(use_graph, end)
            else:
                self._run_per_tick(end)
        if self.verbose >= 1:
            print("Simulation complete.")

    def _run_fused(self, use_graph, end=None):
        end = self.nt if end is None else end
        eng = self._ensure_engine()
        ds = self._instance("DiseaseState_ABM")
        # Seeds are applied inside DiseaseState_ABM.step (t >= 1), between the DS
        # progression and RI: such a day runs as VD+DS, host seeding, RI+SIA+TX+log.
        seed_days = sorted(t for t in (ds.seed_schedule if ds else {}) if max(self.t, 1) <= t < end)
        t = self.t
        for sd in seed_days:
            if t < sd:
                eng.run_days(t, sd, use_graph)
                t = sd
            self.t = sd
            eng.step(sd, engmod.STAGE_VITAL | engmod.STAGE_DISEASE)
            self._mark_device_newer()
            ds._apply_seed_schedule(sd)
            self._push_touched()
            eng.step(sd, engmod.STAGE_RI | engmod.STAGE_SIA | engmod.STAGE_TRANSMISSION | engmod.STAGE_LOG)
            t = sd + 1
        if t < end:
            eng.run_days(t, end, use_graph)
        eng.sync()
        self._mark_device_newer()
        stop = int(eng.stop_days()[0])
        if stop >= 0:
            self.t = stop + 1
            self.should_stop = True
            if self.verbose >= 1:
                print(f"[SEIR_ABM] Early stopping at t={self.t}: no E/I and no future seed_schedule events.")
        else:
            self.t = end

    def _run_per_tick(self, end=None):
        for tick in range(self.t, self.nt if end is None else end):
            if tick == 0:
                self.log_results(0)
                self.t += 1
                continue
            for component in self.instances:
                with self.perf_stats.start(type(component).__name__ + ".step()"):
                    component.step()
            self.log_results(tick)
            self.t += 1
            if self.should_stop:
                break

    def log_results(self, t):
        for component in self.instances:
            component.log(t)

    def plot(self, save=False, results_path=None):
        """Plotting is not part of the GPU step; kept for API compatibility."""
        return


# =============================================================================
class VitalDynamics_ABM:
    @classmethod
    def init_from_file(cls, sim):
        """For a model loaded from a snapshot (model.py:1529-1537): results rows, birth rates
        and the death table; the loaded agents keep their dates."""
        self = cls.__new__(cls)
        self._common_init(sim)
        p, pars = sim.people, sim.pars
        for name in ("date_of_birth", "date_of_death"):
            if not p.has(name):
                p.add_scalar_property(name, dtype=np.int32, default=AGENT_COLUMNS[name][1])
        self.km_cd = create_cumulative_deaths(int(np.sum(pars.init_pop)), max_age_years=100)
        if pars.cbr is not None:
            for name in ("births", "deaths"):
                if not self.results.has(name):
                    self.results.add_array_property(name, shape=(sim.nt, len(self.nodes)), dtype=np.int32)
        self._initialize_birth_rates()
        return self

    """Births (node tails) and deaths by precomputed date of death (model.py:1521-1764)."""

    def __init__(self, sim):
        self._common_init(sim)
        self._initialize_ages_and_births()
        self._initialize_deaths()
        self._initialize_birth_rates()

    def _common_init(self, sim):
        self.sim = sim
        self.people = sim.people
        self.nodes = sim.nodes
        self.results = sim.results
        self.pars = sim.pars
        self.step_size = self.pars.step_size_VitalDynamics_ABM
        self.verbose = sim.pars["verbose"] if "verbose" in sim.pars else 1

    def _sample_ages(self, n):
        pars, rng = self.pars, self.sim.rng
        if pars.age_model is not None:
            kind, mean_y, cap_y = pars.age_model
            if kind != "exponential":
                raise ValueError(f"unknown age_model {kind}")
            ages = np.minimum(rng.exponential(mean_y * 365.0, n), cap_y * 365.0).astype(np.int32)
        elif pars.age_pyramid_path is not None:
            pyr = load_pyramid_csv(pars.age_pyramid_path)
            w = (pyr[:, 2] + pyr[:, 3]).astype(np.float64)
            bins = rng.choice(len(pyr), size=n, p=w / w.sum())
            lo = np.maximum(pyr[:, 0] * 365, 1)
            hi = (pyr[:, 1] + 1) * 365
            ages = (lo[bins] + np.floor(rng.random(n) * (hi[bins] - lo[bins]))).astype(np.int32)
        else:
            ages = np.minimum(rng.exponential(25 * 365.0, n), 80 * 365.0).astype(np.int32)
        ages[ages == 0] = 1
        return ages

    def _initialize_ages_and_births(self):
        p = self.people
        p.add_scalar_property("date_of_birth", dtype=np.int32, default=-1)
        if self.sim.device_init:
            self.sim.init_spec["ages"] = self._age_spec()
            return
        dob = p.raw("date_of_birth")
        table = self.pars.init_sus_by_age
        for n in range(len(self.nodes)):
            sl = p.initial_slice(n)
            if table is not None:
                dob[sl] = -self._sample_ages_by_bin(table[table["node_id"] == n], sl.stop - sl.start)
            else:
                dob[sl] = -self._sample_ages(sl.stop - sl.start)

    def _sample_ages_by_bin(self, rows, n):
        """Ages (days) of a node's susceptibles from its init_sus_by_age rows: a bin drawn in
        proportion to n_susceptible, the day uniform in [age_min_yr, age_max_yr) x 365
        (model.py:1553-1573)."""
        rng = self.sim.rng
        w = rows["n_susceptible"].to_numpy(dtype=np.float64)
        lo = (rows["age_min_yr"].to_numpy(dtype=np.float64) * 365).astype(np.int64)
        hi = (rows["age_max_yr"].to_numpy(dtype=np.float64) * 365).astype(np.int64)
        if n == 0 or w.sum() <= 0:
            return np.ones(n, dtype=np.int32)
        b = rng.choice(len(w), size=n, p=w / w.sum())
        ages = lo[b] + np.floor(rng.random(n) * np.maximum(hi[b] - lo[b], 1)).astype(np.int64)
        ages[ages <= 0] = 1
        return ages.astype(np.int32)

    def _age_spec(self):
        """Age model of the initial population for the device initialisation."""
        pars = self.pars
        if pars.age_model is not None:
            kind, mean_y, cap_y = pars.age_model
            if kind != "exponential":
                raise ValueError(f"unknown age_model {kind}")
            return ("exponential", mean_y * 365.0, cap_y * 365.0)
        if pars.age_pyramid_path is not None:
            pyr = load_pyramid_csv(pars.age_pyramid_path)
            w = (pyr[:, 2] + pyr[:, 3]).astype(np.float64)
            return ("pyramid", np.cumsum(w / w.sum()), np.maximum(pyr[:, 0] * 365, 1), (pyr[:, 1] + 1) * 365)
        return ("exponential", 25 * 365.0, 80 * 365.0)

    def _initialize_deaths(self):
        pars, p = self.pars, self.people
        self.km_cd = create_cumulative_deaths(int(np.sum(pars.init_pop)), max_age_years=100)
        if pars.cbr is not None:
            self.results.add_array_property("births", shape=(self.sim.nt, len(self.nodes)), dtype=np.int32)
            self.results.add_array_property("deaths", shape=(self.sim.nt, len(self.nodes)), dtype=np.int32)
            p.add_scalar_property("date_of_death", dtype=np.int32, default=0)
            if self.sim.device_init:
                self.sim.init_spec["vd"] = True
                return
            dod, dob = p.raw("date_of_death"), p.raw("date_of_birth")
            life = np.zeros(len(self.nodes))
            for n in range(len(self.nodes)):
                sl = p.initial_slice(n)
                ages = -dob[sl].astype(np.int64)
                at_death = km_age_at_death(self.km_cd, ages, self.sim.rng)
                dod[sl] = at_death - ages
                life[n] = at_death.mean() / 365 if len(at_death) else 0.0
            pars.life_expectancies = life
            if pars.init_sus_by_age is not None:
                self._immune_deaths()

    def _immune_deaths(self):
        """Deaths of the immune population that init_sus_by_age leaves out of the frame,
        drawn up front: per node and day Poisson(sum over age bins of n_immune / (remaining
        life in days)), remaining life = max(node life expectancy - mid-bin age, 1) years;
        none on day 0 (model.py:1651-1669)."""
        pars = self.pars
        df = pars.init_sus_by_age
        mid = (df["age_min_yr"].to_numpy(dtype=np.float64) + df["age_max_yr"].to_numpy(dtype=np.float64)) / 2
        nid = df["node_id"].to_numpy(dtype=np.int64)
        remaining = np.maximum(pars.life_expectancies[nid] - mid, 1.0)
        rate = df["n_immune"].to_numpy(dtype=np.float64) / (remaining * 365)
        per_node = np.bincount(nid, weights=rate, minlength=len(self.nodes))
        deaths = self.results.raw("deaths")
        deaths += self.sim.rng.poisson(np.outer(np.ones(deaths.shape[0]), per_node)).astype(np.int32)
        deaths[0, :] = 0

    def _initialize_birth_rates(self):
        pars = self.pars
        self.birth_rate = np.zeros(len(self.nodes))
        if pars.cbr is not None:
            self.birth_rate[:] = np.broadcast_to(np.atleast_1d(np.asarray(pars.cbr, dtype=np.float64)), self.birth_rate.shape) / (365 * 1000)

    def step(self):
        self.sim._engine_step(engmod.STAGE_VITAL)

    def log(self, t):
        pass

    def plot(self, save=False, results_path=None):
        pass


def load_pyramid_csv(path):
    """Age pyramid CSV (columns Age,M,F with bins 'a-b' and 'a+') -> int64 [bins, 4] = min, max, M, F."""
    rows = []
    with open(path) as fh:
        next(fh)
        for line in fh:
            line = line.strip()
            if not line:
                continue
            age, m, f = line.split(",")
            if age.endswith("+"):
                lo = hi = int(age[:-1])
            else:
                lo, hi = (int(x) for x in age.split("-"))
            rows.append((lo, hi, int(m), int(f)))
    return np.array(rows, dtype=np.int64)


def km_age_at_death(cd, ages_days, rng):
    """Age at death (days) conditional on surviving to ``ages_days``: year of death drawn in
    proportion to the deaths of years >= the current year of age in the cumulative table ``cd``,
    day uniform within that year and after the current age."""
    cd = np.asarray(cd, dtype=np.int64)
    a = np.asarray(ages_days, dtype=np.int64)
    max_year = len(cd) - 1
    ay = np.minimum(a // 365, max_year)
    lo = np.where(ay == 0, 0, cd[np.maximum(ay - 1, 0)])
    hi = cd[max_year]
    span = hi - lo
    v = lo + np.floor(rng.random(len(a)) * np.maximum(span, 0)).astype(np.int64)
    yod = np.where(span > 0, np.searchsorted(cd, v, side="right"), ay)
    lo_d = np.maximum(yod * 365, a + 1)
    hi_d = (yod + 1) * 365
    return np.where(hi_d > lo_d, lo_d + np.floor(rng.random(len(a)) * (hi_d - lo_d)).astype(np.int64), lo_d)


# =============================================================================
class DiseaseState_ABM:
    """E->I->R progression, paralysis, seeding, early stop (model.py:505-811)."""

    @classmethod
    def init_from_file(cls, sim):
        """For a model loaded from a snapshot (model.py:506-524): results rows; the slots not
        yet handed out get heterogeneity values and timers drawn, potentially_paralyzed,
        paralyzed and ipv_protected -1."""
        self = cls.__new__(cls)
        self._common_init(sim)
        self._initialize_results_arrays()
        p, pars, rng = sim.people, sim.pars, sim.rng
        for name in AGENT_COLUMNS:
            if name != "ri_timer" and not p.has(name):
                dt, default = AGENT_COLUMNS[name]
                p.add_scalar_property(name, dtype=dt, default=default)
        risk, inf = p.raw("acq_risk_multiplier"), p.raw("daily_infectivity")
        et, it, pt = p.raw("exposure_timer"), p.raw("infection_timer"), p.raw("paralysis_timer")
        for n in range(p.n_nodes):
            a, b = int(p.live_end[n]), int(p.seg_end[n])
            if b <= a:
                continue
            populate_heterogeneous_values(a, b, risk, inf, pars, rng)
            et[a:b] = timer_from_samples(pars.dur_exp(b - a))
            it[a:b] = timer_from_samples(pars.dur_inf(b - a))
            raw = pars.t_to_paralysis(b - a) - et[a:b]
            pt[a:b] = np.clip(raw, 0, np.minimum(it[a:b], 127)).astype(np.int64).astype(np.uint16)
            for name in ("potentially_paralyzed", "paralyzed", "ipv_protected"):
                p.raw(name)[a:b] = -1
        return self

    def __init__(self, sim):
        self._common_init(sim)
        self._initialize_results_arrays()
        pars, p, rng = self.pars, self.people, sim.rng
        cap = p.capacity
        p.add_scalar_property("exposure_timer", dtype=np.uint16, default=0)
        p.add_scalar_property("infection_timer", dtype=np.uint16, default=0)
        p.add_scalar_property("paralysis_timer", dtype=np.uint16, default=0)
        et, it, pt = p.raw("exposure_timer"), p.raw("infection_timer"), p.raw("paralysis_timer")
        B = 1 << 23 if not sim.device_init else 0
        for b0 in range(0, cap if B else 0, B or 1):
            b1 = min(cap, b0 + B)
            et[b0:b1] = timer_from_samples(pars.dur_exp(b1 - b0))
            it[b0:b1] = timer_from_samples(pars.dur_inf(b1 - b0))
        for b0 in range(0, cap if B else 0, B or 1):
            b1 = min(cap, b0 + B)
            raw = pars.t_to_paralysis(b1 - b0) - et[b0:b1]
            pt[b0:b1] = np.clip(raw, 0, np.minimum(it[b0:b1], 127)).astype(np.int64).astype(np.uint16)
        state = p.raw("disease_state")
        N = len(pars.init_pop)
        if pars.init_sus_by_age is not None:
            self._init_susceptibles_only()
        elif isinstance(pars.init_immun, (float, int)) and not isinstance(pars.init_immun, bool):
            fr = np.full(N, float(pars.init_immun))
        elif isinstance(pars.init_immun, (list, tuple, np.ndarray)):
            fr = np.asarray(pars.init_immun, dtype=np.float64)
            if fr.shape != pars.init_pop.shape:
                raise ValueError("init_immun must match init_pop shape")
        else:
            raise ValueError(f"Unsupported init_immun type: {type(pars.init_immun)}")
        for n in range(N if pars.init_sus_by_age is None else 0):
            if not (0.0 <= fr[n] <= 1.0):
                raise ValueError(f"Invalid immunity fraction: {fr[n]} for node {n}")
            k = int(fr[n] * pars.init_pop[n])
            if k > 0 and not sim.device_init:
                sl = p.initial_slice(n)
                state[sl.start + rng.choice(sl.stop - sl.start, size=k, replace=False)] = 3
        if sim.device_init:
            if np.any(fr != fr[0]):
                raise ValueError("device initialisation takes one init_immun fraction for every node")
            sim.init_spec["init_immun"] = float(fr[0])
            if pars.init_recovered_by_age is not None:
                if not p.has("date_of_birth"):
                    raise ValueError("init_recovered_by_age needs VitalDynamics_ABM (ages)")
                sim.init_spec["bands"] = [(lo * 365, hi * 365, frac) for lo, hi, frac in pars.init_recovered_by_age]
        elif pars.init_recovered_by_age is not None and pars.init_sus_by_age is None:
            dob = p.raw("date_of_birth") if p.has("date_of_birth") else None
            if dob is None:
                raise ValueError("init_recovered_by_age needs VitalDynamics_ABM (ages)")
            for n in range(N):
                sl = p.initial_slice(n)
                ages = -dob[sl].astype(np.int64)
                u = rng.random(sl.stop - sl.start)
                rec = np.zeros(sl.stop - sl.start, bool)
                for lo_y, hi_y, frac in pars.init_recovered_by_age:
                    band = (ages >= lo_y * 365) & (ages < hi_y * 365)
                    rec |= band & (u < frac)
                state[sl][rec] = 3
        total = int(p.init_pop.sum())     # initial agents of the frame
        idx = []
        if isinstance(pars.init_prev, float):
            idx = p.initial_to_slot(rng.choice(total, size=int(total * pars.init_prev), replace=False))
        elif isinstance(pars.init_prev, (int, np.integer)) and not isinstance(pars.init_prev, bool):
            idx = p.initial_to_slot(rng.choice(total, size=min(int(pars.init_prev), total), replace=False))
        elif isinstance(pars.init_prev, (list, tuple, np.ndarray)):
            if len(pars.init_prev) != N:
                raise ValueError(f"Length mismatch: init_prev has {len(pars.init_prev)} entries, expected {N} nodes.")
            parts = []
            for n, prev in enumerate(pars.init_prev):
                if not isinstance(prev, numbers.Real):
                    raise ValueError(f"Unsupported value in init_prev list at node {n}: {prev}")
                k = int(pars.init_pop[n] * prev) if 0 < prev < 1 else min(int(prev), int(pars.init_pop[n]))
                sl = p.initial_slice(n)
                # every initial agent is alive here (immune ones included)
                alive = np.arange(sl.start, sl.stop) if sim.device_init else sl.start + np.nonzero(state[sl] >= 0)[0]
                k = min(k, len(alive))
                if k > 0:
                    parts.append(rng.choice(alive, size=k, replace=False))
            idx = np.concatenate(parts) if parts else []
        else:
            raise ValueError(f"Unsupported init_prev type: {type(pars.init_prev)}")
        if sim.device_init:
            sim.init_spec["infected"] = np.asarray(idx, dtype=np.int64)
        elif len(idx):
            state[np.asarray(idx, dtype=np.int64)] = 2

    def _init_susceptibles_only(self):
        """init_sus_by_age (model.py:614-665): every agent of the frame is susceptible; the
        immune remainder of each node (init_pop - init_sus) is added to every row of
        results.R, less the cumulative immune deaths VitalDynamics drew up front; per age
        bin, n_ipv_protected of the node's agents in the bin are IPV-protected."""
        pars, p, rng = self.pars, self.people, self.sim.rng
        R = self.results.raw("R")
        R += (pars.init_pop - pars.init_sus).astype(np.int32)[None, :]
        df = pars.init_sus_by_age
        ipv = p.raw("ipv_protected")
        if "n_ipv_protected" in df.columns and p.has("date_of_birth"):
            dob = p.raw("date_of_birth")
            rows = df[df["n_ipv_protected"] > 0]
            for node in rows["node_id"].unique():
                sl = p.initial_slice(int(node))
                ages_y = -dob[sl].astype(np.float64) / 365.0
                for _, row in rows[rows["node_id"] == node].iterrows():
                    elig = sl.start + np.nonzero((ages_y >= row["age_min_yr"]) & (ages_y < row["age_max_yr"]))[0]
                    k = min(int(row["n_ipv_protected"]), len(elig))
                    if k > 0:
                        ipv[rng.choice(elig, size=k, replace=False)] = 1
        if self.results.has("deaths"):
            R -= np.cumsum(self.results.raw("deaths"), axis=0).astype(np.int32)

    def _common_init(self, sim):
        self.sim = sim
        self.people = sim.people
        self.pars = sim.pars
        self.nodes = sim.nodes
        self.results = sim.results
        self.verbose = self.pars["verbose"] if "verbose" in self.pars else 1
        self.seed_schedule = defaultdict(list)
        if self.pars.seed_schedule is not None:
            lookup = self.pars.node_lookup or {}
            for entry in self.pars.seed_schedule:
                if "date" in entry and "dot_name" in entry:
                    t = (to_date(entry["date"]) - to_date(self.pars.start_date)).days
                    nid = next((k for k, v in lookup.items() if v["dot_name"] == entry["dot_name"]), None)
                    if nid is not None:
                        self.seed_schedule[t].append((nid, entry["prevalence"]))
                elif "timestep" in entry and "node_id" in entry:
                    self.seed_schedule[entry["timestep"]].append((entry["node_id"], entry["prevalence"]))

    def _initialize_results_arrays(self):
        nt, N, K = self.sim.nt, len(self.nodes), len(self.sim.pars.strain_ids)
        for name in ("S", "E", "I", "R"):
            self.results.add_array_property(name, shape=(nt, N), dtype=np.int32)
        self.results.add_array_property("E_by_strain", shape=(nt, N, K), dtype=np.int32)
        self.results.add_array_property("I_by_strain", shape=(nt, N, K), dtype=np.int32)
        for name in ("potentially_paralyzed", "paralyzed", "new_potentially_paralyzed", "new_paralyzed", "pop"):
            self.results.add_array_property(name, shape=(nt, N), dtype=np.int32)
        self.results.raw("pop")[0] = self.sim.pars.init_pop

    def _apply_seed_schedule(self, t):
        """Seed infections listed for tick t: n agents alive in the node, chosen uniformly
        without replacement, become I; timers <= 0 are redrawn (model.py:759-786)."""
        p, sim = self.people, self.sim
        for node_id, value in self.seed_schedule.get(t, []):
            sl = p.node_slice(node_id)
            state = p.disease_state
            cand = sl.start + np.nonzero(state[sl] >= 0)[0]
            if isinstance(value, float):
                n_seed = int(len(cand) * value)
            elif isinstance(value, (int, np.integer)):
                n_seed = min(int(value), len(cand))
            else:
                raise ValueError(f"Unsupported seed value type: {type(value)}")
            if n_seed <= 0:
                continue
            sel = sim.rng.choice(cand, size=n_seed, replace=False)
            state[sel] = 2
            it = p.infection_timer
            zero = sel[it[sel] == 0]
            it[zero] = timer_from_samples(self.pars.dur_inf(len(zero)))
            if self.verbose >= 1:
                print(f"[DiseaseState_ABM] t={t}: Seeded {n_seed} infections in node {node_id}")

    def step(self):
        sim = self.sim
        sim._engine_step(engmod.STAGE_DISEASE)
        if sim.t in self.seed_schedule:
            self._apply_seed_schedule(sim.t)
        if self.pars["stop_if_no_cases"]:
            stop = int(sim._engine.stop_days()[0])
            if stop >= 0 and stop <= sim.t:
                sim.should_stop = True

    def log(self, t):
        pass

    def plot(self, save=False, results_path=None):
        pass


# =============================================================================
def populate_heterogeneous_values(start, end, acq_risk_out, infectivity_out, pars, rng):
    """Correlated per-agent acquisition risk and daily infectivity (model.py:816-866).

    "reference" model: risk ~ lognormal(mean 1, variance risk_mult_var), infectivity ~
    Gamma(shape 1, mean infectivity_mean) joined by a Gaussian copula with normal correlation
    2 sin(pi corr / 6).  "lognormal" model: both are mean-1 lognormals with standard deviation
    heterogeneity_sigma (infectivity scaled by infectivity_mean), same copula.
    """
    mean_inf = pars.infectivity_mean
    if mean_inf is None:
        mean_inf = pars.r0 / np.mean(pars.dur_inf(1000))
    rho = 2.0 * np.sin(np.pi * pars.corr_risk_inf / 6)
    L = np.linalg.cholesky(np.array([[1.0, rho], [rho, 1.0]]))
    if pars.heterogeneity_model == "reference":
        var_ln = pars.risk_mult_var
        mu_r = np.log(1.0 / np.sqrt(var_ln + 1.0))
        sg_r = np.sqrt(np.log(var_ln + 1.0))
    elif pars.heterogeneity_model == "lognormal":
        s = float(pars.heterogeneity_sigma)
        mu_r = np.log(1.0 / np.sqrt(s * s + 1.0))
        sg_r = np.sqrt(np.log(s * s + 1.0))
    else:
        raise ValueError(f"unknown heterogeneity_model {pars.heterogeneity_model}")
    from scipy.special import log_ndtr

    B = 1 << 22
    for b0 in range(start, end, B):
        b1 = min(end, b0 + B)
        if not pars.individual_heterogeneity:
            acq_risk_out[b0:b1] = 1.0
            infectivity_out[b0:b1] = mean_inf
            continue
        z = rng.standard_normal((b1 - b0, 2)) @ L.T
        acq_risk_out[b0:b1] = np.exp(mu_r + sg_r * z[:, 0])
        if pars.heterogeneity_model == "reference":
            # Gamma(1, scale).ppf(Phi(z)) = -scale * log(1 - Phi(z)) = -scale * log Phi(-z)
            infectivity_out[b0:b1] = -max(mean_inf, 1e-10) * log_ndtr(-z[:, 1])
        else:
            infectivity_out[b0:b1] = mean_inf * np.exp(mu_r + sg_r * z[:, 1])


def heterogeneity_spec(pars):
    """(model, log-mean, log-sd, infectivity mean, normal correlation) of the copula of
    populate_heterogeneous_values, for the device initialisation: model 0 constants, 1 both
    lognormal, 2 lognormal risk with gamma(1) infectivity."""
    mean_inf = pars.infectivity_mean
    if mean_inf is None:
        mean_inf = pars.r0 / np.mean(pars.dur_inf(1000))
    if not pars.individual_heterogeneity:
        return (0, 0.0, 0.0, float(mean_inf), 0.0)
    rho = 2.0 * np.sin(np.pi * pars.corr_risk_inf / 6)
    if pars.heterogeneity_model == "reference":
        var_ln = pars.risk_mult_var
        return (2, np.log(1.0 / np.sqrt(var_ln + 1.0)), np.sqrt(np.log(var_ln + 1.0)), float(max(mean_inf, 1e-10)), rho)
    if pars.heterogeneity_model == "lognormal":
        sd = float(pars.heterogeneity_sigma)
        return (1, np.log(1.0 / np.sqrt(sd * sd + 1.0)), np.sqrt(np.log(sd * sd + 1.0)), float(mean_inf), rho)
    raise ValueError(f"unknown heterogeneity_model {pars.heterogeneity_model}")


class Transmission_ABM:
    """Force of infection over the node network and infection draws (model.py:1152-1490)."""

    @classmethod
    def init_from_file(cls, sim):
        """For a model loaded from a snapshot (model.py:1169-1188): network and results; when
        pars.r0 differs from the snapshot's (pars.old_r0), every daily_infectivity is scaled
        by r0 / old_r0."""
        self = cls.__new__(cls)
        self.sim = sim
        self.people = sim.people
        self.nodes = np.arange(len(sim.pars.init_pop))
        self.pars = sim.pars
        self.results = sim.results
        self.verbose = sim.pars["verbose"] if "verbose" in sim.pars else 1
        self.r0_scalars = np.array(self.pars.r0_scalars)
        old = getattr(sim.pars, "old_r0", None)
        if old is not None and sim.pars.r0 != old:
            inf = sim.people.raw("daily_infectivity")
            inf *= np.float32(sim.pars.r0 / old)
        self._initialize_common()
        return self

    def __init__(self, sim):
        self.sim = sim
        self.people = sim.people
        self.nodes = np.arange(len(sim.pars.init_pop))
        self.pars = sim.pars
        self.results = sim.results
        self.verbose = sim.pars["verbose"] if "verbose" in sim.pars else 1
        self.r0_scalars = np.array(self.pars.r0_scalars)
        self._initialize_people_fields()
        self._initialize_common()

    def _initialize_people_fields(self):
        p = self.people
        p.add_scalar_property("acq_risk_multiplier", dtype=np.float32, default=1.0)
        p.add_scalar_property("daily_infectivity", dtype=np.float32, default=1.0)
        if self.sim.device_init:
            self.sim.init_spec["het"] = heterogeneity_spec(self.pars)
            return
        populate_heterogeneous_values(0, p.capacity, p.raw("acq_risk_multiplier"), p.raw("daily_infectivity"), self.pars, self.sim.rng)

    def _initialize_common(self):
        pars = self.pars
        init_pops = np.asarray(pars.init_pop, dtype=np.float64)
        if pars.distances is not None:
            dist = np.asarray(pars.distances, dtype=np.float64)
        else:
            lookup = pars.node_lookup
            ids = sorted(lookup.keys())
            lats = np.array([lookup[i]["lat"] for i in ids])
            lons = np.array([lookup[i]["lon"] for i in ids])
            dist = distance(lats[:, None], lons[:, None], lats[None, :], lons[None, :])
            off = ~np.eye(len(ids), dtype=bool)
            dist[off & (dist == 0)] = 1.0
        method = pars.migration_method.lower()
        self.distances = dist
        self.gravity_spec = None
        if method == "gravity":
            k = pars.gravity_k * 10 ** (pars.gravity_k_exponent)
            self.gravity_spec = (float(k), float(pars.gravity_a), float(pars.gravity_b), float(pars.gravity_c),
                                 float(pars.max_migr_frac))
            net = gravity(init_pops, dist, k, pars.gravity_a, pars.gravity_b, pars.gravity_c)
            net /= np.power(init_pops.sum(), pars.gravity_c)
        elif method == "radiation":
            net = radiation(init_pops, dist, 10**pars.radiation_k_log10, include_home=False)
        else:
            raise ValueError(f"Unknown migration method: {pars.migration_method}")
        self.network = row_normalizer(net, pars.max_migr_frac)
        nt, N, K = self.sim.nt, len(self.nodes), len(pars.strain_ids)
        self.results.add_array_property("new_exposed", shape=(nt, N), dtype=np.int32)
        self.results.add_array_property("new_exposed_by_strain", shape=(nt, N, K), dtype=np.int32)

    def step(self):
        self.sim._engine_step(engmod.STAGE_TRANSMISSION)

    def log(self, t):
        self.sim._engine_step(engmod.STAGE_LOG)

    def plot(self, save=False, results_path=""):
        pass


# =============================================================================
class RI_ABM:
    """Routine OPV + IPV every step_size days (model.py:1858-1991)."""

    @classmethod
    def init_from_file(cls, sim):
        """For a model loaded from a snapshot (model.py:1873-1885): results rows; the loaded
        agents keep their ri_timer."""
        self = cls.__new__(cls)
        self.sim = sim
        self.step_size = sim.pars.step_size_RI_ABM
        self.people = sim.people
        self.nodes = sim.nodes
        self.pars = sim.pars
        self.results = sim.results
        self.verbose = self.pars["verbose"] if "verbose" in self.pars else 1
        if not self.people.has("ri_timer"):
            self.people.add_scalar_property("ri_timer", dtype=np.int16, default=-1)
        self._initialize_results()
        return self

    def __init__(self, sim):
        self.sim = sim
        self.step_size = sim.pars.step_size_RI_ABM
        self.people = sim.people
        self.nodes = sim.nodes
        self.pars = sim.pars
        self.results = sim.results
        self.verbose = self.pars["verbose"] if "verbose" in self.pars else 1
        p = self.people
        p.add_scalar_property("ri_timer", dtype=np.int16, default=-1)
        ri, lo_hi = p.raw("ri_timer"), self.pars.ri_age_days
        dob = p.raw("date_of_birth") if p.has("date_of_birth") else None
        if sim.device_init:
            if dob is None:
                raise ValueError("device initialisation of RI timers needs VitalDynamics_ABM (ages)")
            sim.init_spec["ri"] = (float(lo_hi[0]), float(lo_hi[1]))
        for n in range(len(self.nodes) if not sim.device_init else 0):
            sl = p.initial_slice(n)
            base = dob[sl] if dob is not None else np.full(sl.stop - sl.start, -1, np.int32)
            days = sim.rng.uniform(lo_hi[0], lo_hi[1], sl.stop - sl.start) if lo_hi[1] > lo_hi[0] else np.full(sl.stop - sl.start, float(lo_hi[0]))
            ri[sl] = (base + days).astype(np.int32).astype(np.int16)
        self._initialize_results()

    def _initialize_results(self):
        nt, N, K = self.sim.nt, len(self.sim.nodes), len(self.pars.strain_ids)
        self.results.add_array_property("ri_vaccinated", shape=(nt, N), dtype=np.int32)
        self.results.add_array_property("ri_protected", shape=(nt, N), dtype=np.int32)
        self.results.add_array_property("ri_new_exposed_by_strain", shape=(nt, N, K), dtype=np.int32)
        self.results.add_array_property("ipv_vaccinated", shape=(nt, N), dtype=np.int32)

    def step(self):
        self.sim._engine_step(engmod.STAGE_RI)

    def log(self, t):
        pass

    def plot(self, save=False, results_path=None):
        pass


class SIA_ABM:
    """Campaign vaccination on scheduled dates (model.py:2063-2162)."""

    @classmethod
    def init_from_file(cls, sim):
        """For a model loaded from a snapshot (model.py:2070-2075): results and schedule."""
        return cls(sim)

    def __init__(self, sim):
        self.sim = sim
        self.people = sim.people
        self.nodes = sim.nodes
        self.pars = sim.pars
        self.results = sim.results
        self.verbose = sim.pars["verbose"] if "verbose" in sim.pars else 1
        nt, N, K = sim.nt, len(self.nodes), len(self.pars.strain_ids)
        self.results.add_array_property("sia_vaccinated", shape=(nt, N), dtype=np.int32)
        self.results.add_array_property("sia_protected", shape=(nt, N), dtype=np.int32)
        self.results.add_array_property("sia_new_exposed_by_strain", shape=(nt, N, K), dtype=np.int32)
        sched = self.pars["sia_schedule"] if "sia_schedule" in self.pars else None
        self.sia_schedule = [] if sched is None else sched
        for event in self.sia_schedule:
            event["date"] = to_date(event["date"])

    def step(self):
        self.sim._engine_step(engmod.STAGE_SIA)

    def log(self, t):
        pass

    def plot(self, save=False, results_path=None):
        pass
