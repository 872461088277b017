"""Run a constructed SEIR_ABM on the HIP path or on the CPU oracle, with one view of the outcome."""

import copy

import numpy as np
import pytest

from oracle.oracle import OracleSim

RESULT_NAMES = [
    "S", "E", "I", "R", "E_by_strain", "I_by_strain", "potentially_paralyzed", "paralyzed",
    "new_potentially_paralyzed", "new_paralyzed", "pop", "births", "deaths", "new_exposed",
    "new_exposed_by_strain", "ri_vaccinated", "ri_protected", "ri_new_exposed_by_strain", "ipv_vaccinated",
    "sia_vaccinated", "sia_protected", "sia_new_exposed_by_strain",
]

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


class Outcome:
    def __init__(self, res, cols, live_end, seg_begin, ticks):
        self.res = res
        self.cols = cols
        self.live_end = live_end
        self.seg_begin = seg_begin
        self.ticks = ticks

    def born_slots(self):
        return np.concatenate([np.arange(b, e) for b, e in zip(self.seg_begin, self.live_end, strict=True)])


def _sampler(dist):
    """(rng, n) -> n samples of a duration Distribution drawn from `rng`, with the numpy
    Generator calls the product's distributions make (constant, gamma, normal, poisson)."""
    d, p = dist.dist_type, dist.pars
    if d == "constant":
        return lambda rng, n: np.full(n, p.get("value", 1))
    if d == "gamma":
        return lambda rng, n: rng.gamma(p.get("shape", 2.0), p.get("scale", 1.0), n)
    if d == "normal":
        return lambda rng, n: rng.normal(p.get("mean", 0.0), p.get("std", 1.0), n)
    if d == "poisson":
        return lambda rng, n: rng.poisson(p.get("lam", 5), n)
    raise NotImplementedError(d)


def seeding_of(sim):
    """The oracle's seed_schedule inputs for a constructed (not yet run) sim: the raw entries,
    node lookup, start date, a copy of the host Generator in its current state (the product
    draws its seeds from sim.rng during the run) and the infectious-duration sampler."""
    if sim.pars.seed_schedule is None:
        return None
    return {"seed_schedule": sim.pars.seed_schedule, "node_lookup": sim.pars.node_lookup,
            "start_date": sim.pars.start_date, "rng": copy.deepcopy(sim.rng), "dur_inf": _sampler(sim.pars.dur_inf)}


def oracle_of(sim, **kw):
    """OracleSim of a constructed (not yet run) sim."""
    results0 = {k: sim.results.raw(k) for k in ("R", "deaths") if sim.results.has(k)}
    return OracleSim(sim.export_tables(), pop0=sim.pars.init_pop, results0=results0, seeding=seeding_of(sim), **kw)


def run(sim, backend):
    """Run the simulation to completion on `backend` ("oracle" or "gpu")."""
    if backend == "oracle":
        o = oracle_of(sim).run()
        return Outcome(o.res, o.cols, o.live_end, np.asarray(sim.people.seg_begin), o.ticks_run)
    sim.run()
    res = {k: np.array(getattr(sim.results, k)) for k in RESULT_NAMES if sim.results.has(k)}
    cols = {k: np.array(getattr(sim.people, k)) for k in sim.people.column_names()}
    return Outcome(res, cols, np.array(sim.people.live_end), np.asarray(sim.people.seg_begin), sim.t)
