"""Host-side logic and the C-ABI library, without a GPU.

No entry point that launches work is called here: the library is loaded, its
exported symbols are checked against include/lp_step.h, the pure queries (ABI
version, struct size, bind names) are compared with the host mirror, and
lp_create is called only with configurations it rejects before any HIP call.
"""

import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

import laser_polio_amd as lp
from laser_polio_amd import engine as em
from laser_polio_amd import model as mdl
from laser_polio_amd.pars import default_run_order

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lp_step.h")
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def _declared():
    names = set()
    for h in HEADERS:
        names |= set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(lp_\w+)\(", open(h).read(), re.M))
    return sorted(names)


def _lib():
    if not os.path.exists(em.LIB_PATH):
        pytest.skip("liblpstep.so not built (run __graft_entry__.build())")
    return em.load_library()


def test_header_declares_the_expected_api():
    names = _declared()
    for f in ("lp_create", "lp_bind", "lp_prepare", "lp_run_days", "lp_step_stage", "lp_step_vital",
              "lp_step_disease", "lp_step_ri", "lp_step_sia", "lp_step_transmission", "lp_log", "lp_sync",
              "lp_stop_days", "lp_destroy", "lp_last_error"):
        assert f in names
    for f in ("lp_mf_days", "lp_mf_adjoint_days"):
        assert f in names


def test_library_exports_every_declared_symbol():
    lib = _lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", em.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (lp_\w+)$", out, re.M))
    assert set(_declared()) <= exported
    # no entry point of the boundary is exported C++-mangled
    mangled = [s for s in re.findall(r"\b[TD] (\S+)$", out, re.M) if s.startswith("_Z")]
    assert not [m for m in mangled for name in _declared() if name in m]


def test_abi_version_and_struct_layout():
    lib = _lib()
    assert lib.lp_abi_version() == 1
    assert lib.lp_sizeof_sim_params() == em.SIM_PARAMS_DTYPE.itemsize == 88
    text = open(HEADER).read()
    for name, val in (("LP_TILE", 1024), ("LP_CHUNK", 256), ("LP_MAX_STRAINS", em.MAX_STRAINS)):
        assert re.search(rf"#define {name} {val}\b", text), name


def test_create_rejects_out_of_range_configs():
    """lp_create checks its limits before any HIP call: capacity below 2^32 and a multiple
    of LP_TILE, at most 32767 nodes, 1-4 strains, 1 <= nt <= 2^23 (list entries hold event
    days in 24 bits)."""
    lib = _lib()
    good = dict(capacity=1 << 20, n_nodes=4, n_sims=1, n_strains=1, nt=366, n_sia_events=0, mix_mode=0)
    bad = [dict(nt=(1 << 23) + 1), dict(nt=0), dict(capacity=1 << 32), dict(capacity=1000), dict(n_nodes=32768),
           dict(n_strains=5), dict(n_strains=0), dict(n_sims=0), dict(mix_mode=2), dict(mix_mode=-1), dict(ei_grid=32),
           dict(sample_grid=70000), dict(commit_grid=-1)]
    for over in bad:
        cfg = em.LPConfig(**{**good, **over})
        ctx = ctypes.c_void_p()
        assert lib.lp_create(ctypes.byref(cfg), ctypes.byref(ctx)) == -1, over   # LP_ERR_ARG
        assert not ctx.value


def test_bind_names_cover_host_columns_and_results():
    names = set(_lib().lp_bind_names().decode().split(","))
    assert set(mdl.AGENT_COLUMNS) <= names
    assert {b for b, *_ in mdl.RESULT_ARRAYS.values()} <= names
    for tbl in ("seg_begin", "live_end", "seg_end", "network_t", "km_cd", "day_flags", "season", "sim_params"):
        assert tbl in names


def test_product_path_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    sim = lp.SEIR_ABM(lp.PropertySet({"dur": 3, "init_pop": np.array([100, 100]), "verbose": 0, "seed": 1}))
    sim.components = [lp.DiseaseState_ABM, lp.Transmission_ABM]
    with pytest.raises(em.EngineUnavailable):
        sim.run()


def test_components_follow_default_run_order():
    sim = lp.SEIR_ABM(lp.PropertySet({"dur": 2, "init_pop": np.array([50, 60]), "verbose": 0, "seed": 2}))
    sim.components = [lp.Transmission_ABM, lp.SIA_ABM, lp.VitalDynamics_ABM, lp.RI_ABM, lp.DiseaseState_ABM]
    assert [c.__name__ for c in sim.components] == list(default_run_order)
    assert [type(i).__name__ for i in sim.instances] == list(default_run_order)


def test_propertyset_merge_semantics():
    ps = lp.PropertySet({"a": 1, "b": 2})
    ps <<= {"a": 5}
    assert ps.a == 5 and ps["b"] == 2
    with pytest.raises(Exception):
        ps <<= {"zzz": 1}
    ps += {"c": 3}
    assert ps.c == 3 and len(ps) == 3
    assert (ps | {"a": 9, "d": 4}).to_dict() == {"a": 9, "b": 2, "c": 3, "d": 4}


def test_node_segmented_layout():
    pars = {"dur": 30, "init_pop": np.array([3000, 10, 1, 2500]), "cbr": np.array([40, 30, 20, 35]), "verbose": 0, "seed": 4,
            "init_immun": 0.5, "init_prev": 0.01, "r0_scalars": np.ones(4), "distances": np.ones((4, 4))}
    sim = lp.SEIR_ABM(lp.PropertySet(pars))
    sim.components = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.Transmission_ABM]
    tb = sim.export_tables()
    sb, le, se = tb.seg_begin, tb.live_end, tb.seg_end
    assert sb[0] == 0 and np.all(sb[1:] == se[:-1]) and se[-1] == tb.capacity
    assert np.all(sb % 1024 == 0) and np.all(se % 1024 == 0)
    assert np.array_equal(le - sb, pars["init_pop"])
    st, node = tb.cols["disease_state"], tb.cols["node_id"]
    for n in range(4):
        assert np.all(st[sb[n] : le[n]] >= 0) and np.all(node[sb[n] : le[n]] == n)
        assert np.all(st[le[n] : se[n]] == -1)
    # room for the expected births of the run (calc_capacity plus margin)
    births = pars["init_pop"] * pars["cbr"] / 1000 * 30 / 365
    assert np.all(se - le >= births)
    slots = sim.people.initial_to_slot(np.arange(int(pars["init_pop"].sum())))
    assert np.array_equal(np.sort(slots), np.concatenate([np.arange(sb[n], le[n]) for n in range(4)]))


def test_day_flags_schedule():
    pars = {"dur": 30, "init_pop": np.array([500, 500]), "cbr": np.array([30, 30]), "verbose": 0, "seed": 5, "init_immun": 0.5, "init_prev": 0.01,
            "r0_scalars": np.ones(2), "distances": np.array([[0, 1], [1, 0]]), "vx_prob_sia": [0.7, 0.7], "vx_prob_ri": 0.5,
            "sia_schedule": [{"date": lp.date("2019-01-10"), "nodes": [0], "age_range": (0, 1825), "vaccinetype": "nOPV2"}],
            "start_date": lp.date("2019-01-01")}
    sim = lp.SEIR_ABM(lp.PropertySet(pars))
    sim.components = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.RI_ABM, lp.SIA_ABM, lp.Transmission_ABM]
    tb = sim.export_tables()
    f = tb.day_flags
    # day 0 only logs the initial counts
    assert f[0] == em.DAY_TX
    assert all(f[t] & em.DAY_TX and f[t] & em.DAY_DS for t in range(1, sim.nt))
    vd_days = [t for t in range(1, sim.nt) if f[t] & em.DAY_VD_STEP]
    vd, ri = sim._instance("VitalDynamics_ABM"), sim._instance("RI_ABM")
    assert vd.step_size == 7 and ri.step_size == 14
    assert vd_days == [t for t in range(1, sim.nt) if t % 7 == 0]
    assert all(f[t] & em.DAY_VD_COPY for t in range(1, sim.nt) if t % 7)
    sia_days = [t for t in range(1, sim.nt) if f[t] & em.DAY_SIA]
    assert sia_days == [9]
    ri_days = [t for t in range(1, sim.nt) if f[t] & em.DAY_RI]
    assert ri_days == [t for t in range(1, sim.nt) if t % 14 == 0]


def test_thresholds_round_trip():
    p = np.array([0.0, 1e-9, 0.25, 0.999999, 1.0])
    thr = em.thr_from_prob64(p)
    r = np.array([0, 1, 2**30, 2**32 - 1], np.uint64)
    # r32 < thr(p)  <=>  r32 * 2^-32 < p
    for pi, ti in zip(p, thr):
        assert np.array_equal(r < ti, r.astype(np.float64) * 2.0**-32 < pi)


def test_autotune_keeps_the_default_within_the_margin():
    """A launch-shape candidate replaces the default only when its trial is more than
    TUNE_MARGIN faster: short trials cannot separate closer candidates."""
    pick = lp.BatchedSims._pick
    m = lp.BatchedSims.TUNE_MARGIN
    assert pick({0: 1.00, 1024: 0.99, 2048: 1.02}, 0) == 0
    assert pick({0: 1.00, 1024: 1.0 - 2 * m, 2048: 1.02}, 0) == 1024
    assert pick({4: 3.1, 8: 3.0, 16: 3.05}, 16) == 16
    assert pick({4: 3.1, 8: 2.8, 16: 3.05}, 16) == 8
    assert pick({4: 3.1, 8: 3.0}, 16) == 8          # no default among the candidates
