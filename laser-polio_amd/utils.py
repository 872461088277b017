"""Date, seasonality, mortality-table and timing helpers.

Behaviour follows src/laser_polio/utils.py: ``date``/``daterange``
(utils.py:137-161), ``get_doy``/``get_seasonality`` (utils.py:589-625),
``create_cumulative_deaths`` (utils.py:795-816) and ``TimingStats``
(utils.py:819-945, reduced to what SEIR_ABM uses).  ``calc_capacity`` restates
laser_core.utils.calc_capacity (not vendored by the reference): population
grown at the daily crude birth rate for ``nticks`` days.
"""

import calendar
import datetime as dt
import time
from contextlib import contextmanager

import numpy as np


def date(value):
    if isinstance(value, dt.datetime):
        return value.date()
    if isinstance(value, dt.date):
        return value
    return dt.datetime.strptime(value, "%Y-%m-%d").date()


def daterange(start_date, days):
    start = date(start_date)
    return np.array([start + dt.timedelta(days=i) for i in range(days)])


def get_doy(sim):
    """Day of year (1-based) of ``sim.datevec[sim.t]``."""
    return sim.datevec[sim.t].timetuple().tm_yday


def seasonality_of(day, amplitude, peak_doy):
    """1 + amplitude * cos(2 pi (doy - peak) / days_in_year), doy 1-based, leap aware."""
    doy = day.timetuple().tm_yday
    days_in_year = 366 if calendar.isleap(day.year) else 365
    return 1 + amplitude * np.cos(2 * np.pi * (doy - peak_doy) / days_in_year)


def get_seasonality(sim):
    return seasonality_of(sim.datevec[sim.t], sim.pars["seasonal_amplitude"], sim.pars["seasonal_peak_doy"])


def create_cumulative_deaths(total_population, max_age_years):
    """Cumulative deaths by year of age with mortality 1e-4 * 2^(age/10) per person-year."""
    ages_years = np.arange(max_age_years + 1)
    mortality_rates = 0.0001 * (2 ** (ages_years / 10))
    return np.cumsum(mortality_rates * total_population).astype(np.int64)


def calc_capacity(population, nticks, cbr):
    """Population after ``nticks`` days of growth at crude birth rate ``cbr`` (per 1000 per year)."""
    daily = (float(cbr) / 1000.0) / 365.0
    return int(np.ceil(float(population) * (1.0 + daily) ** int(nticks)))


class TimingStats:
    """Accumulated wall-clock nanoseconds per label."""

    def __init__(self):
        self.stats = {}

    @contextmanager
    def start(self, label):
        t0 = time.perf_counter_ns()
        try:
            yield
        finally:
            self.stats[label] = self.stats.get(label, 0) + time.perf_counter_ns() - t0

    def log(self, logger=None):
        if logger is None:
            return
        for k, v in self.stats.items():
            logger.info(f"{k}: {v / 1e6:.3f} ms")
