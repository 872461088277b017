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


# -- results files (utils.py:690-786, 1040-1160) --------------------------------------
def add_temporal_groupings(df, time_config):
    """A `time_period` column: dates binned at time_config["bins"] (left-closed), labelled
    time_config["labels"] (utils.py:1040-1073)."""
    import pandas as pd

    df = df.copy()
    if "bins" in time_config and "labels" in time_config:
        edges = [pd.Timestamp.min, *[pd.Timestamp(d) for d in time_config["bins"]], pd.Timestamp.max]
        df["time_period"] = pd.cut(df["date"], bins=edges, labels=time_config["labels"], right=False)
    return df


def add_regional_groupings(df, region_groupings=None, grouping_level="adm0", regions_yaml_path=None):
    """adm0 / adm1 / adm01 columns split from dot_name ("AFRO:COUNTRY:STATE:..."), and a
    `region` column at `grouping_level`; countries in `region_groupings` take the custom
    regions of the regions YAML file (country -> region -> patterns), others keep
    grouping_level (utils.py:1076-1160)."""
    import pandas as pd

    df = df.copy()
    parts = df["dot_name"].astype(str).str.split(":", expand=True)
    df["adm0"] = parts[1] if 1 in parts.columns else None
    df["adm1"] = parts[2] if 2 in parts.columns else None
    df["adm01"] = df["adm0"] + ":" + df["adm1"]
    if grouping_level not in {"adm0", "adm01", "dot_name"}:
        raise ValueError(f"Invalid grouping_level: {grouping_level}. Must be one of 'adm0', 'adm01', or 'dot_name'.")
    df["region"] = df[grouping_level]
    if region_groupings:
        import yaml

        try:
            with open(regions_yaml_path) as fh:
                regions = yaml.safe_load(fh)
        except (FileNotFoundError, TypeError):
            print(f"Warning: Could not load {regions_yaml_path}, using {grouping_level} for all countries")
            return df
        for country in region_groupings:
            groups = regions.get(country.upper())
            if groups is None:
                print(f"Warning: Custom regions not found for {country} in regions.yaml, using adm0")
                continue
            in_country = df["adm0"] == country.upper()
            for name, patterns in groups.items():
                hit = pd.Series(False, index=df.index)
                for pat in patterns:
                    if grouping_level == "adm01" and ":" in pat:
                        hit |= df["adm01"] == pat
                    else:
                        hit |= df["dot_name"].astype(str).str.contains(pat, case=False, na=False, regex=False)
                df.loc[in_country & hit, "region"] = name
    return df


def save_sim_results(sim, filename="simulation_results.csv", summary_config=None):
    """One row per (timestep, node) with the date, dot_name and the S, E, I, R, P, births,
    deaths, new_exposed, potentially_paralyzed, new_potentially_paralyzed and new_paralyzed
    results; with `summary_config`, time_periods / region_groupings / grouping_level
    columns.  Written as CSV, or Parquet for a .parquet name; HDF5 needs PyTables, which
    this environment lacks, so a .h5 name raises (utils.py:690-786).  Returns the frame."""
    from pathlib import Path

    import pandas as pd

    nt, N, r = sim.nt, len(sim.nodes), sim.results
    lookup = sim.pars.node_lookup or {}

    def col(name):
        return np.asarray(getattr(r, name)).reshape(-1) if r.has(name) else np.zeros(nt * N, np.int32)

    data = {
        "timestep": np.repeat(np.arange(nt), N),
        "date": np.repeat(np.asarray(sim.datevec), N),
        "node": np.tile(np.arange(N), nt),
        "dot_name": np.tile(np.array([lookup.get(n, {}).get("dot_name", "UNKNOWN") for n in range(N)], dtype=object), nt),
        "S": col("S"), "E": col("E"), "I": col("I"), "R": col("R"), "P": col("paralyzed"),
        "births": col("births"), "deaths": col("deaths"), "new_exposed": col("new_exposed"),
        "potentially_paralyzed": col("potentially_paralyzed"),
        "new_potentially_paralyzed": col("new_potentially_paralyzed"), "new_paralyzed": col("new_paralyzed"),
    }
    df = pd.DataFrame(data)
    df["dot_name"] = df["dot_name"].astype("category")
    if summary_config is not None:
        df["date"] = pd.to_datetime(df["date"])
        if "time_periods" in summary_config:
            df = add_temporal_groupings(df, summary_config["time_periods"])
        groups = summary_config.get("region_groupings")
        level = summary_config.get("grouping_level")
        df = add_regional_groupings(df, groups, level or ("adm0" if groups is not None else "dot_name"),
                                    summary_config.get("regions_yaml_path"))
    suffix = Path(filename).suffix
    if suffix == ".h5":
        raise RuntimeError("HDF5 results need PyTables, which is not installed here; use a .csv or .parquet name")
    if suffix == ".parquet":
        df.to_parquet(filename, index=False)
    else:
        df.to_csv(filename, index=False)
    return df
