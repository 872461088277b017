"""Run a constructed SEIR_ABM on the HIP path or on the CPU oracle, with one view of the outcome."""

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


def run(sim, backend):
    """Run the simulation to completion on `backend` ("oracle" or "gpu")."""
    if backend == "oracle":
        results0 = {k: sim.results.raw(k) for k in ("R",) if sim.results.has(k)}
        o = OracleSim(sim.export_tables(), pop0=sim.pars.init_pop, results0=results0).run()
        return Outcome(o.res, o.cols, o.live_end, np.asarray(sim.people.seg_begin), o.ticks_run)
    sim.run()
    res = {k: np.array(getattr(sim.results, k)) for k in RESULT_NAMES if sim.results.has(k)}
    cols = {k: np.array(getattr(sim.people, k)) for k in sim.people.column_names()}
    return Outcome(res, cols, np.array(sim.people.live_end), np.asarray(sim.people.seg_begin), sim.t)
