"""Default parameters and component run order.

Keys and values follow src/laser_polio/pars.py:10-99.  Additions used by the
synthetic workloads (all default to the reference behaviour):
  age_model               None -> ages from ``age_pyramid_path``; ("exponential", mean_y, cap_y)
  init_recovered_by_age   list of (min_age_y, max_age_y, fraction) applied per node after init_immun
  heterogeneity_model     "reference" (lognormal risk, gamma infectivity) or "lognormal"
  heterogeneity_sigma     std of the mean-1 lognormal multipliers for "lognormal"
  infectivity_mean        mean daily infectivity; None -> r0 / mean(dur_inf(1000)) as the reference
  ri_age_days             (lo, hi) days from birth to RI for the initial population (reference 42..98)
  ri_newborn_timer        ri_timer of newborns (reference 182)
  tiles_per_block         agent-pass tiles per workgroup (0 = autotuned default)
  init_on_device          draw the per-agent columns on the GPU (lp_init_population) instead of on
                          the host: same distributions, different numbers; initial immunity and
                          chronically-missed agents become per-agent Bernoulli draws
  mix_mode                0: network mix in the oracle's fixed order (bit-identical replays);
                          1: the prevalence x network contraction on MFMA (float32 within 1e-6
                          of float64, not replay-identical)
  ei_grid, sample_grid    workgroups of the E / I list pass and the round-1 sampler (0 = the
                          library's defaults; each 64..65535)
"""

import datetime

import numpy as np

from . import distributions as lpd
from .propertyset import PropertySet

__all__ = ["default_pars", "default_run_order"]

default_pars = PropertySet(
    {
        "seed": None,
        "start_date": datetime.date(2019, 1, 1),
        "dur": 30,
        "init_pop": [15000, 10000],
        "init_immun": [0.0, 0.0],
        "init_sus_by_age": None,
        "age_pyramid_path": None,
        "cbr": [37, 41],
        "strain_ids": {"VDPV2": 0, "Sabin2": 1, "nOPV2": 2},
        "strain_r0_scalars": {0: 1.0, 1: 0.25, 2: 0.125},
        "init_prev": [0.0, 0.0],
        "seed_schedule": None,
        "r0": 14,
        "r0_scalars": [0.8, 1.2],
        "seasonal_amplitude": 0.125,
        "seasonal_peak_doy": 180,
        "individual_heterogeneity": True,
        "risk_mult_var": 4.0,
        "corr_risk_inf": 0.8,
        "dur_exp": lpd.poisson(lam=3),
        "dur_inf": lpd.gamma(shape=4.51, scale=5.32),
        "t_to_paralysis": lpd.lognormal(mean=12.5, sigma=3.5),
        "p_paralysis": 1 / 2000,
        "shp": None,
        "node_lookup": None,
        "distances": np.array([[0, 100], [100, 0]]),
        "node_seeding_dispersion": 1000,
        "node_seeding_zero_inflation": 0.0,
        "migration_method": "radiation",
        "radiation_k_log10": -0.3,
        "gravity_k": 1.0,
        "gravity_k_exponent": 0.0,
        "gravity_a": 1,
        "gravity_b": 1,
        "gravity_c": 2.0,
        "max_migr_frac": 0.1,
        "vx_prob_ri": None,
        "vx_prob_ipv": None,
        "ipv_start_year": 2015,
        "sia_schedule": None,
        "vx_prob_sia": None,
        "missed_frac": 0.0,
        "vx_efficacy": {
            "perfect": 1.0,
            "bOPV": 0,
            "f-IPV": 0,
            "IPV": 0,
            "IPV + bOPV": 0,
            "mOPV2": 0.7,
            "nOPV2": 0.7 * 0.8,
            "nOPV2 + fIPV": 0.7 * 0.8,
            "topv": 0.5,
        },
        "step_size_VitalDynamics_ABM": 7,
        "step_size_DiseaseState_ABM": 1,
        "step_size_RI_ABM": 14,
        "step_size_SIA_ABM": 1,
        "step_size_Transmission_ABM": 1,
        "actual_data": None,
        "summary_config": None,
        "verbose": 1,
        "stop_if_no_cases": True,
        # additions (see module docstring)
        "age_model": None,
        "init_recovered_by_age": None,
        "heterogeneity_model": "reference",
        "heterogeneity_sigma": 0.5,
        "infectivity_mean": None,
        "ri_age_days": (42, 98),
        "ri_newborn_timer": 182,
        "tiles_per_block": 0,
        "init_on_device": False,
        "mix_mode": 0,
        "ei_grid": 0,
        "sample_grid": 0,
    }
)

default_run_order = ["VitalDynamics_ABM", "DiseaseState_ABM", "RI_ABM", "SIA_ABM", "Transmission_ABM"]
