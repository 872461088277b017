"""Small simulation cases shared by the CPU and GPU tests."""

import numpy as np

import laser_polio_amd as lp

ALL = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.RI_ABM, lp.SIA_ABM, lp.Transmission_ABM]


def make(case, seed=7, **over):
    """Build (sim, components) for a named case; sims are constructed but not run."""
    base = {
        "seed": seed,
        "verbose": 0,
        "stop_if_no_cases": False,
        "migration_method": "gravity",
        "gravity_k": 0.5,
        "max_migr_frac": 0.3,
    }
    comps = ALL
    if case == "small3":
        p = dict(dur=40, init_pop=np.array([30000, 20000, 15000]), cbr=np.array([35.0] * 3), r0=14,
                 r0_scalars=np.array([1.0, 0.8, 1.2]), init_immun=[0.6] * 3, init_prev=[20, 0, 0],
                 distances=np.array([[0, 50, 90], [50, 0, 60], [90, 60, 0]]), vx_prob_ri=0.5)
        comps = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.RI_ABM, lp.Transmission_ABM]
    elif case == "sia":
        p = dict(dur=60, init_pop=np.array([20000, 12000, 9000, 5000]), cbr=np.array([40.0]), r0=10,
                 r0_scalars=np.array([1.0, 1.0, 0.7, 1.3]), init_immun=0.5, init_prev=[10, 0, 0, 0],
                 distances=np.array([[0, 30, 60, 90], [30, 0, 30, 60], [60, 30, 0, 30], [90, 60, 30, 0]]),
                 vx_prob_ri=0.7, vx_prob_ipv=0.4, missed_frac=0.1, ri_newborn_timer=42, ri_age_days=(42, 42),
                 vx_prob_sia=[0.8, 0.6, 0.9, 0.5],
                 sia_schedule=[{"date": "2019-01-11", "nodes": [0, 2], "age_range": (0, 5 * 365), "vaccinetype": "nOPV2"},
                               {"date": "2019-01-11", "nodes": [1, 2, 3], "age_range": (0, 15 * 365), "vaccinetype": "topv"},
                               {"date": "2019-02-05", "nodes": [3], "age_range": (100, 2000), "vaccinetype": "mOPV2"}],
                 step_size_RI_ABM=14)
    elif case == "c0mini":
        rng = np.random.default_rng(42)
        n = 14
        pop = np.round(rng.lognormal(np.log(20000), 0.5, n)).astype(int)
        d = rng.uniform(20, 250, (n, n))
        d = np.triu(d, 1)
        d = d + d.T
        p = dict(dur=120, init_pop=pop, cbr=np.array([35.0]), r0=14, r0_scalars=rng.lognormal(0, 0.3, n),
                 seasonal_amplitude=0.4, seasonal_peak_doy=180, init_immun=0.0,
                 init_recovered_by_age=[(15, 200, 1.0), (0, 5, 0.8)], age_model=("exponential", 25, 80),
                 init_prev=[0.005] + [0.0] * (n - 1), distances=d, dur_exp=lp.normal(mean=3, std=1),
                 heterogeneity_model="lognormal", heterogeneity_sigma=0.5, infectivity_mean=14 / 23.9932)
        comps = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.Transmission_ABM]
    elif case == "hot":
        # huge R0: most susceptibles exposed in a day -> many sampling rounds and take-all nodes
        p = dict(dur=12, init_pop=np.array([3000, 800, 200]), cbr=np.array([30.0]), r0=400,
                 r0_scalars=np.array([1.0, 1.5, 2.0]), init_immun=0.2, init_prev=[30, 10, 2],
                 distances=np.array([[0, 1, 2], [1, 0, 1], [2, 1, 0]]), dur_exp=lp.constant(2), dur_inf=lp.constant(3))
        comps = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.Transmission_ABM]
    elif case == "hotbig":
        # huge R0 on large nodes: later sampling rounds of more probes than k_finish keeps
        # in LDS (its selected-list path)
        p = dict(dur=6, init_pop=np.array([60000, 30000]), cbr=np.array([30.0]), r0=400,
                 r0_scalars=np.array([1.0, 1.5]), init_immun=0.1, init_prev=[400, 100],
                 distances=np.array([[0, 1], [1, 0]]), dur_exp=lp.constant(2), dur_inf=lp.constant(3))
        comps = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.Transmission_ABM]
    elif case == "bignode":
        # a node of more than 64 spans (over 1M slots): the samplers' span search in
        # global memory instead of one span per lane
        p = dict(dur=10, init_pop=np.array([1_150_000, 40000]), cbr=np.array([35.0]), r0=25,
                 r0_scalars=np.array([1.0, 1.0]), init_immun=0.3, init_prev=[3000, 20],
                 distances=np.array([[0, 20], [20, 0]]))
        comps = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.Transmission_ABM]
    elif case == "zinb":
        # infection only in node 0; other nodes get importations through the zero-inflated NB branch
        p = dict(dur=80, init_pop=np.array([40000, 30000, 30000, 20000]), cbr=None, r0=12,
                 r0_scalars=np.array([1.0, 1.0, 1.0, 1.0]), init_immun=0.3, init_prev=[100, 0, 0, 0],
                 distances=np.array([[0, 10, 20, 30], [10, 0, 10, 20], [20, 10, 0, 10], [30, 20, 10, 0]]),
                 gravity_k=2000.0, node_seeding_zero_inflation=0.3, node_seeding_dispersion=3)
        comps = [lp.DiseaseState_ABM, lp.Transmission_ABM]
    elif case == "stop":
        # epidemic dies out -> early stop
        p = dict(dur=200, init_pop=np.array([2000, 1500]), cbr=np.array([20.0]), r0=0.3,
                 r0_scalars=np.array([1.0, 1.0]), init_immun=0.0, init_prev=[5, 0], individual_heterogeneity=False,
                 dur_inf=lp.constant(4), distances=np.array([[0, 10], [10, 0]]), stop_if_no_cases=True)
        comps = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.Transmission_ABM]
    elif case == "campaigns":
        # SIA on a plain day (t=10), on a VD step day (t=7) and on a VD + RI day (t=14):
        # every combination of births/deaths, RI and SIA in one day
        p = dict(dur=45, init_pop=np.array([9000, 7000, 5000]), cbr=np.array([45.0]), r0=9,
                 r0_scalars=np.array([1.0, 1.2, 0.9]), init_immun=0.4, init_prev=[15, 3, 0],
                 distances=np.array([[0, 40, 70], [40, 0, 40], [70, 40, 0]]), vx_prob_ri=0.8, vx_prob_ipv=0.3,
                 missed_frac=0.05, ri_newborn_timer=14, ri_age_days=(14, 28), vx_prob_sia=[0.7, 0.9, 0.6],
                 sia_schedule=[{"date": "2019-01-08", "nodes": [0, 1], "age_range": (0, 5 * 365), "vaccinetype": "nOPV2"},
                               {"date": "2019-01-11", "nodes": [2], "age_range": (0, 10 * 365), "vaccinetype": "topv"},
                               {"date": "2019-01-15", "nodes": [0, 1, 2], "age_range": (0, 3 * 365), "vaccinetype": "mOPV2"}],
                 step_size_RI_ABM=14)
    elif case == "ri_only":
        # RI and SIA without VitalDynamics: RI-only days and an SIA on an RI day
        p = dict(dur=40, init_pop=np.array([8000, 6000]), cbr=None, r0=8, r0_scalars=np.array([1.0, 1.1]),
                 init_immun=0.3, init_prev=[10, 0], distances=np.array([[0, 30], [30, 0]]), vx_prob_ri=0.9,
                 vx_prob_ipv=0.5, ri_age_days=(0, 60), vx_prob_sia=[0.8, 0.8],
                 sia_schedule=[{"date": "2019-01-15", "nodes": [0, 1], "age_range": (0, 5 * 365), "vaccinetype": "nOPV2"},
                               {"date": "2019-01-20", "nodes": [1], "age_range": (0, 5 * 365), "vaccinetype": "topv"}],
                 step_size_RI_ABM=14)
        comps = [lp.DiseaseState_ABM, lp.RI_ABM, lp.SIA_ABM, lp.Transmission_ABM]
    elif case == "many":
        # more nodes than the per-workgroup LDS tables hold (8200 > 8191): the list pass
        # and the sampler look nodes up in global memory and round 1 is planned by k_plan
        rng = np.random.default_rng(3)
        n = 8200
        xy = rng.uniform(0, 1000, (n, 2))
        d = np.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1))
        p = dict(dur=9, init_pop=rng.integers(20, 60, n), cbr=np.array([40.0]), r0=30,
                 r0_scalars=rng.lognormal(0, 0.3, n), init_immun=0.2, init_prev=[5] * 50 + [0] * (n - 50),
                 distances=d)
        comps = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.Transmission_ABM]
    elif case == "seeds":
        # seed_schedule injections on a plain day (float prevalence by date + dot_name), on
        # day 2 (int count by timestep), on a VD step day (t=7), on a VD + RI day (t=14,
        # more than the node's live agents) and on an SIA day (t=10): each is a
        # VD + DS stage call, host seeding, then a RECOUNT stage call of RI + SIA + TX + log
        lookup = {i: {"dot_name": f"AFRO:TESTLAND:REGION:N{i}", "lat": 0.0, "lon": float(i)} for i in range(4)}
        p = dict(dur=30, init_pop=np.array([9000, 6000, 4000, 2500]), cbr=np.array([40.0]), r0=6,
                 r0_scalars=np.array([1.0, 1.1, 0.9, 1.2]), init_immun=0.5, init_prev=[10, 0, 0, 0],
                 distances=np.array([[0, 40, 70, 90], [40, 0, 40, 60], [70, 40, 0, 30], [90, 60, 30, 0]]),
                 vx_prob_ri=0.6, vx_prob_ipv=0.3, ri_newborn_timer=14, ri_age_days=(14, 28), vx_prob_sia=[0.7] * 4,
                 sia_schedule=[{"date": "2019-01-11", "nodes": [1, 3], "age_range": (0, 5 * 365), "vaccinetype": "nOPV2"}],
                 node_lookup=lookup,
                 seed_schedule=[{"date": "2019-01-02", "dot_name": "AFRO:TESTLAND:REGION:N2", "prevalence": 0.02},
                                {"timestep": 2, "node_id": 3, "prevalence": 15},
                                {"timestep": 7, "node_id": 1, "prevalence": 0.05},
                                {"timestep": 10, "node_id": 0, "prevalence": 40},
                                {"timestep": 14, "node_id": 3, "prevalence": 10**6},
                                {"timestep": 14, "node_id": 2, "prevalence": 0.01}])
    elif case == "stopseed":
        # the epidemic dies out and the early-stop rule waits for seeds still ahead (t=60
        # and t=62; a seed on the day of the check itself does not count as ahead,
        # model.py:792), the seeds re-ignite transmission, and the run stops after it dies out again
        p = dict(dur=200, init_pop=np.array([3000, 2000]), cbr=np.array([20.0]), r0=0.5,
                 r0_scalars=np.array([1.0, 1.0]), init_immun=0.0, init_prev=[5, 0], dur_inf=lp.constant(4),
                 distances=np.array([[0, 10], [10, 0]]), stop_if_no_cases=True,
                 seed_schedule=[{"timestep": 60, "node_id": 1, "prevalence": 20},
                                {"timestep": 62, "node_id": 0, "prevalence": 0.01}])
        comps = [lp.VitalDynamics_ABM, lp.DiseaseState_ABM, lp.Transmission_ABM]
    else:
        raise KeyError(case)
    base.update(p)
    base.update(over)
    sim = lp.SEIR_ABM(lp.PropertySet(base))
    sim.components = comps
    return sim


CASES = ["small3", "sia", "c0mini", "hot", "hotbig", "bignode", "zinb", "stop", "campaigns", "ri_only"]
