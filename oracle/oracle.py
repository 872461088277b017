"""Driver of the C oracle (lp_oracle.cpp).  TEST INFRASTRUCTURE ONLY.

Used by tests/ as the checker of the HIP path, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg; nothing in the product imports it.

``OracleSim(tables)`` runs the reference day loop (SEIR_ABM.run,
src/laser_polio/model.py:246-289) over the oracle's restated kernels, in the
reference's component order and pass structure:

  tick 0                 Transmission_ABM.log -> count_SEIRP          model.py:1461-1483
  VitalDynamics_ABM.step get_deaths, births, pop bookkeeping          model.py:1689-1757
  DiseaseState_ABM.step  disease_state_step_kernel, early-stop rule   model.py:715-795
  RI_ABM.step            fast_ri + result rows                        model.py:1920-1985
  SIA_ABM.step           fast_sia per event + result rows             model.py:2099-2156
  Transmission_ABM.step  tx_step_prep, node map + count draws,
                         tx_infect_nb                                 model.py:1281-1441
  Transmission_ABM.log   count_SEIRP                                  model.py:1461-1483

Parity status of the oracle's functions (how each is pinned, see tests/):
  philox                 Random123 known-answer vectors (tests/test_oracle_math.py)
  orc_expm1f/expf/log1pf float64 numpy within 2 ulp over sweeps
  disease_state_step     reference KATs test_progression_without_transmission,
                         test_disease_timers_with_trans_explicit,
                         test_paralysis_progression_manual; brute force
  get_deaths            bookkeeping KATs of tests/test_vital_dynamics.py
  births                 bookkeeping KATs; expected births computed in float64
                         exactly as model.py:1712-1716 and checked against numpy's
                         evaluation of that expression over a sweep of rates and
                         populations (tests/test_oracle_math.py)
  seed_schedule          reference test_seed_schedule (tests/test_diseasestate_abm.py:
                         398-457) ported with synthetic nodes; seeded counts and
                         timers checked against the definition
  fast_ri / fast_sia     test_ri_no_effect_on_non_susceptibles, RI windows of
                         test_ri_zero, SIA date/node/age targeting; binomial rates
  tx_step_prep / count   brute-force numpy bincount
  node map (FOI)         float64 closed form (oracle/fp64.py) to 1e-6
  count draws            Poisson / NB moments and pmf (chi-square) on sweeps
  tx_infect rounds       inclusion probabilities vs exact enumeration on tiny nodes
  KM age at death        parity unpinned (laser_core KaplanMeierEstimator is not
                         vendored by the reference); distribution checked against
                         the cumulative-deaths table it is defined from
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liblporacle.so")

DAY_VD_STEP, DAY_VD_COPY, DAY_DS, DAY_RI, DAY_SIA, DAY_TX, DAY_STOPCHK = (1, 2, 4, 8, 16, 32, 64)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_PTR_FIELDS = ("state", "strain", "node_id", "exp_timer", "inf_timer", "par_timer", "pp", "paralyzed", "ipv",
               "missed", "risk", "infect", "dob", "dod", "ri_timer", "seg_begin", "live_end", "seg_end", "r0s",
               "keep", "birth_rate_step", "NT", "thr_ri", "thr_ipv", "km_cd")


class OrcModel(ctypes.Structure):
    _fields_ = [
        ("capacity", ctypes.c_int64),
        ("n_nodes", ctypes.c_int32),
        ("n_strains", ctypes.c_int32),
        ("n_threads", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        *[(n, ctypes.c_void_p) for n in _PTR_FIELDS],
        ("km_len", ctypes.c_int32),
        ("newborn_ri_timer", ctypes.c_int32),
        ("strain_r0", ctypes.c_float * 4),
        ("thr_par", ctypes.c_uint64),
        ("thr_zi", ctypes.c_uint64),
        ("disp_r", ctypes.c_int32),
        ("ri_strain", ctypes.c_int32),
        ("ri_step", ctypes.c_int32),
        ("zi", ctypes.c_float),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i32, i64, f32, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_uint64
        L.orc_struct_size.restype = i64
        L.orc_max_threads.restype = i32
        L.orc_philox.argtypes = [vp, u64, vp]
        L.orc_expm1f_c.argtypes = [f32]
        L.orc_expm1f_c.restype = f32
        L.orc_expf_c.argtypes = [f32]
        L.orc_expf_c.restype = f32
        L.orc_log1pf_c.argtypes = [f32]
        L.orc_log1pf_c.restype = f32
        L.orc_poisson.argtypes = [f32, u64, ctypes.c_uint32, i32]
        L.orc_poisson.restype = i64
        L.orc_negbin.argtypes = [i32, f32, u64, ctypes.c_uint32, i32]
        L.orc_negbin.restype = i64
        L.orc_km_age_at_death.argtypes = [vp, i32, u64, u64, i32, i32]
        L.orc_km_age_at_death.restype = i32
        L.orc_birth_expected.argtypes = [ctypes.c_double, i32, vp]
        L.orc_birth_expected.restype = i32
        M = ctypes.POINTER(OrcModel)
        L.orc_get_deaths.argtypes = [M, i32, vp]
        L.orc_births.argtypes = [M, i32, vp, vp]
        L.orc_births.restype = i32
        L.orc_disease_state_step.argtypes = [M, i32, vp, vp]
        L.orc_fast_ri.argtypes = [M, i32, vp, vp, vp]
        L.orc_fast_sia.argtypes = [M, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp]
        L.orc_tx_step_prep.argtypes = [M, vp, vp, vp, vp]
        L.orc_node_map.argtypes = [M, i32, f32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_tx_infect.argtypes = [M, i32, vp, vp, vp, vp, vp]
        L.orc_count_seirp.argtypes = [M, vp, vp, vp, vp, vp, vp]
        if L.orc_struct_size() != ctypes.sizeof(OrcModel):
            raise RuntimeError("orc_model layout mismatch between lp_oracle.cpp and oracle.py")
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def philox(ctr, seed):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox(_p(c), ctypes.c_uint64(seed), _p(out))
    return out


RESULT_SHAPES = {
    "S": 0, "E": 0, "I": 0, "R": 0, "E_by_strain": 1, "I_by_strain": 1, "potentially_paralyzed": 0,
    "paralyzed": 0, "new_potentially_paralyzed": 0, "new_paralyzed": 0, "pop": 0, "births": 0, "deaths": 0,
    "new_exposed": 0, "new_exposed_by_strain": 1, "ri_vaccinated": 0, "ri_protected": 0,
    "ri_new_exposed_by_strain": 1, "ipv_vaccinated": 0, "sia_vaccinated": 0, "sia_protected": 0,
    "sia_new_exposed_by_strain": 1,
}


def parse_seed_schedule(seeding):
    """tick -> [(node, value)] from seed_schedule entries (model.py:535-545): an entry with
    "date" and "dot_name" seeds the node whose node_lookup dot_name matches, on the tick
    (date - start_date).days (entries naming no node are dropped); one with "timestep"
    and "node_id" is taken as it is."""
    import datetime

    out = {}
    lookup = seeding.get("node_lookup") or {}
    start = seeding["start_date"]
    for entry in seeding.get("seed_schedule") or []:
        if "date" in entry and "dot_name" in entry:
            d = entry["date"]
            if isinstance(d, str):
                d = datetime.date.fromisoformat(d)
            node = next((k for k, v in lookup.items() if v["dot_name"] == entry["dot_name"]), None)
            if node is not None:
                out.setdefault((d - start).days, []).append((node, entry["prevalence"]))
        elif "timestep" in entry and "node_id" in entry:
            out.setdefault(entry["timestep"], []).append((entry["node_id"], entry["prevalence"]))
    return out


class OracleSim:
    """One simulation on the CPU oracle, from the product's exported SimTables."""

    def __init__(self, tables, pop0, n_threads=0, results0=None, seeding=None):
        """seeding: None, or a dict with the raw ``seed_schedule`` entries, ``node_lookup``,
        ``start_date``, the numpy Generator ``rng`` the seeds are drawn from and
        ``dur_inf(rng, n)``, the infectious-duration sampler (see _seed)."""
        tb = tables
        self.tb = tb
        self.cols = {k: np.array(v, copy=True) for k, v in tb.cols.items()}
        self.live_end = np.array(tb.live_end, dtype=np.int64, copy=True)
        self.N, self.K, self.nt = tb.n_nodes, tb.n_strains, tb.nt
        self.res = {}
        for k, per in RESULT_SHAPES.items():
            shape = (self.nt, self.N, self.K) if per else (self.nt, self.N)
            self.res[k] = np.zeros(shape, np.int32)
        self.res["pop"][0] = pop0
        if results0:
            for k, v in results0.items():
                self.res[k][...] = v
        self.stop_day = None
        self.keep_alive = []
        m = OrcModel()
        m.capacity = tb.capacity
        m.n_nodes = self.N
        m.n_strains = self.K
        m.n_threads = n_threads
        m.seed = tb.seed
        c = self.cols
        names = {"state": "disease_state", "strain": "strain", "node_id": "node_id", "exp_timer": "exposure_timer",
                 "inf_timer": "infection_timer", "par_timer": "paralysis_timer", "pp": "potentially_paralyzed",
                 "paralyzed": "paralyzed", "ipv": "ipv_protected", "missed": "chronically_missed",
                 "risk": "acq_risk_multiplier", "infect": "daily_infectivity", "dob": "date_of_birth",
                 "dod": "date_of_death", "ri_timer": "ri_timer"}
        for f, col in names.items():
            setattr(m, f, c[col].ctypes.data)
        self.seg_begin = np.ascontiguousarray(tb.seg_begin, dtype=np.int64)
        self.seg_end = np.ascontiguousarray(tb.seg_end, dtype=np.int64)
        self.tables = {
            "r0s": np.ascontiguousarray(tb.r0s, np.float32),
            "keep": np.ascontiguousarray(tb.keep, np.float32),
            "brs": np.ascontiguousarray(tb.birth_rate_step, np.float64),
            "NT": np.ascontiguousarray(tb.network_t, np.float32),
            "thr_ri": np.ascontiguousarray(tb.thr_ri, np.uint64),
            "thr_ipv": np.ascontiguousarray(tb.thr_ipv, np.uint64),
            "km": np.ascontiguousarray(tb.km_cd, np.int64),
        }
        T = self.tables
        m.seg_begin, m.live_end, m.seg_end = self.seg_begin.ctypes.data, self.live_end.ctypes.data, self.seg_end.ctypes.data
        m.r0s, m.keep, m.birth_rate_step, m.NT = T["r0s"].ctypes.data, T["keep"].ctypes.data, T["brs"].ctypes.data, T["NT"].ctypes.data
        m.thr_ri, m.thr_ipv, m.km_cd = T["thr_ri"].ctypes.data, T["thr_ipv"].ctypes.data, T["km"].ctypes.data
        m.km_len = len(T["km"])
        m.newborn_ri_timer = tb.newborn_ri_timer
        for k in range(4):
            m.strain_r0[k] = float(tb.strain_r0[k])
        m.thr_par = tb.thr_paralysis
        m.thr_zi = tb.thr_zi
        m.disp_r = tb.disp_r
        m.ri_strain = tb.ri_strain
        m.ri_step = tb.ri_step
        m.zi = float(tb.zi)
        self.m = m
        self.seeding = seeding
        self.seed_schedule = parse_seed_schedule(seeding) if seeding else {}
        self.node_map_log = {}
        self.record_node_map = False

    # ---------------------------------------------------------------------------
    def run(self, t0=0, t1=None):
        t1 = self.nt if t1 is None else t1
        for t in range(t0, t1):
            self.day(t)
        return self

    def day(self, t):
        L, m, N, K, r = lib(), ctypes.byref(self.m), self.N, self.K, self.res
        f = int(self.tb.day_flags[t])
        if t == 0:
            if f & DAY_TX:
                self._count(0)
            return
        if self.stop_day is not None and t > self.stop_day:
            return
        if (f & DAY_STOPCHK) and self.stop_day is None:
            anyc = int(r["E"][t - 1].sum()) + int(r["I"][t - 1].sum())
            if anyc == 0 and not (self.tb.last_seed_day > t):
                self.stop_day = t
        if f & DAY_VD_STEP:
            deaths = np.zeros(N, np.int32)
            births = np.zeros(N, np.int32)
            L.orc_get_deaths(m, t, _p(deaths))
            pop_prev = np.ascontiguousarray(r["pop"][t - 1])
            if L.orc_births(m, t, _p(pop_prev), _p(births)) != 0:
                raise RuntimeError("oracle: node capacity exhausted")
            r["births"][t] = births
            r["deaths"][t] = deaths
            r["pop"][t] = r["pop"][t - 1] + births - deaths
        elif f & DAY_VD_COPY:
            r["pop"][t] = r["pop"][t - 1]
        if f & DAY_DS:
            a, b = np.zeros(N, np.int32), np.zeros(N, np.int32)
            L.orc_disease_state_step(m, t, _p(a), _p(b))
            r["new_potentially_paralyzed"][t] = a
            r["new_paralyzed"][t] = b
        if (f & DAY_DS) and t in self.seed_schedule:
            self._seed(t)
        if f & DAY_RI:
            a, b, c = (np.zeros(N, np.int32) for _ in range(3))
            L.orc_fast_ri(m, t, _p(a), _p(b), _p(c))
            s = self.tb.ri_strain
            r["ri_vaccinated"][t] = a
            r["ri_protected"][t] = b
            r["new_exposed"][t] += b
            r["new_exposed_by_strain"][t, :, s] += b
            r["ri_new_exposed_by_strain"][t, :, s] = b
            r["ipv_vaccinated"][t] = c
        if f & DAY_SIA:
            b0, n = int(self.tb.day_sia_begin[t]), int(self.tb.day_sia_count[t])
            thr = np.asarray(self.tb.sia_thr, np.uint64).reshape(-1, N, 2)
            for e in range(n):
                ev = b0 + e
                t1 = np.ascontiguousarray(thr[ev, :, 0])
                t2 = np.ascontiguousarray(thr[ev, :, 1])
                tgt = (t1 > 0).astype(np.uint8)
                a, b = np.zeros(N, np.int32), np.zeros(N, np.int32)
                s = int(self.tb.sia_strain[ev])
                L.orc_fast_sia(m, t, e, s, int(self.tb.sia_min_age[ev]), int(self.tb.sia_max_age[ev]),
                               _p(t1), _p(t2), _p(tgt), _p(a), _p(b))
                r["sia_vaccinated"][t] = a
                r["sia_protected"][t] = b
                r["new_exposed"][t] += b
                r["new_exposed_by_strain"][t, :, s] += b
                r["sia_new_exposed_by_strain"][t, :, s] += b
        if f & DAY_TX:
            self._transmission(t)
            self._count(t)

    def _seed(self, t):
        """seed_schedule injections at the end of DiseaseState_ABM.step (model.py:759-786).
        Candidates are the node's live agents in slot order (the reference's index order
        within the node); a float value seeds int(len * value) of them, an int
        min(value, len); they are chosen uniformly without replacement, become I whatever
        their state, and those whose infection timer is 0 get a fresh infectious
        duration, truncated to an integer and clipped to [0, 127] as at initialisation
        (model.py:575-579)."""
        state, itimer = self.cols["disease_state"], self.cols["infection_timer"]
        rng, dur_inf = self.seeding["rng"], self.seeding["dur_inf"]
        for node, value in self.seed_schedule[t]:
            lo, hi = int(self.seg_begin[node]), int(self.live_end[node])
            cand = lo + np.nonzero(state[lo:hi] >= 0)[0]
            if isinstance(value, float):
                n_seed = int(len(cand) * value)
            elif isinstance(value, (int, np.integer)):
                n_seed = min(int(value), len(cand))
            else:
                raise ValueError(f"Unsupported seed value type: {type(value)}")
            if n_seed <= 0:
                continue
            sel = rng.choice(cand, size=n_seed, replace=False)
            state[sel] = 2
            zero = sel[itimer[sel] == 0]
            itimer[zero] = np.clip(np.trunc(np.asarray(dur_inf(rng, len(zero)), np.float64)), 0, 127).astype(np.uint16)

    def _transmission(self, t):
        L, m, N, K = lib(), ctypes.byref(self.m), self.N, self.K
        S_cnt, S_pos = np.zeros(N, np.int32), np.zeros(N, np.int32)
        SW, IB = np.zeros(N, np.int64), np.zeros(N * K, np.int64)
        L.orc_tx_step_prep(m, _p(S_cnt), _p(S_pos), _p(SW), _p(IB))
        p, ex = np.zeros(N * K, np.float32), np.zeros(N, np.float32)
        total, size, take_all = np.zeros(N, np.int64), np.zeros(N, np.int32), np.zeros(N, np.int32)
        thr, bm = np.zeros(N * K, np.uint64), np.zeros(N * K, np.float32)
        pop_t = np.ascontiguousarray(self.res["pop"][t])
        season = float(self.tb.season[t])
        L.orc_node_map(m, t, season, _p(pop_t), _p(S_cnt), _p(S_pos), _p(SW), _p(IB), _p(p), _p(ex),
                       _p(total), _p(size), _p(take_all), _p(thr), _p(bm))
        if self.record_node_map:
            self.node_map_log[t] = {"p": p.reshape(N, K).copy(), "expected": ex.copy(), "total": total.copy(),
                                    "size": size.copy(), "beta_mixed": bm.reshape(N, K).copy(), "SW": SW.copy(),
                                    "IB": IB.reshape(N, K).copy(), "S": S_cnt.copy(), "pop": pop_t.copy()}
        ne = np.zeros(N * K, np.int32)
        L.orc_tx_infect(m, t, _p(S_cnt), _p(size), _p(take_all), _p(thr), _p(ne))
        ne = ne.reshape(N, K)
        self.res["new_exposed"][t] += ne.sum(axis=1).astype(np.int32)
        self.res["new_exposed_by_strain"][t] += ne

    def _count(self, t):
        L, m, N, K = lib(), ctypes.byref(self.m), self.N, self.K
        S, R, PP, P = (np.zeros(N, np.int32) for _ in range(4))
        Eb, Ib = np.zeros(N * K, np.int32), np.zeros(N * K, np.int32)
        L.orc_count_seirp(m, _p(S), _p(Eb), _p(Ib), _p(R), _p(PP), _p(P))
        Eb, Ib = Eb.reshape(N, K), Ib.reshape(N, K)
        r = self.res
        r["S"][t] = S
        r["E"][t] = Eb.sum(axis=1)
        r["I"][t] = Ib.sum(axis=1)
        r["E_by_strain"][t] = Eb
        r["I_by_strain"][t] = Ib
        r["R"][t] += R
        r["potentially_paralyzed"][t] = PP
        r["paralyzed"][t] = P

    @property
    def ticks_run(self):
        return self.nt if self.stop_day is None else self.stop_day + 1
