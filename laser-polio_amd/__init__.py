"""laser-polio daily agent step on MI355X (gfx950).

Host package mirroring the component API of laser-polio
(InstituteforDiseaseModeling/laser-polio, src/laser_polio/model.py): build a
``SEIR_ABM`` from a ``PropertySet``, assign ``components``, ``run()`` and read
``results``.  Every simulated day runs as hand-written HIP kernels in
liblpstep.so (csrc/), bound through the C ABI of include/lp_step.h.
"""

from . import distributions
from .batch import BatchedSims
from .distributions import constant, exponential, gamma, lognormal, lognormal_int, normal, poisson, uniform
from .laserframe import LaserFrame, PeopleFrame
from .model import RI_ABM, SEIR_ABM, SIA_ABM, DiseaseState_ABM, SimTables, Transmission_ABM, VitalDynamics_ABM, load_sim
from .network import distance, gravity, radiation, row_normalizer
from .pars import default_pars, default_run_order
from .propertyset import PropertySet
from .utils import (
    TimingStats,
    add_regional_groupings,
    add_temporal_groupings,
    calc_capacity,
    create_cumulative_deaths,
    date,
    daterange,
    get_doy,
    get_seasonality,
    save_sim_results,
)

__version__ = "0.1.0"
