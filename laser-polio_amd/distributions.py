"""Duration / count distributions used for agent timers.

Same names, parameters and ``dist(size)`` call style as the reference's
src/laser_polio/distributions.py:15-135.  Samples come from the module-level
numpy Generator installed by :func:`set_seed` (SEIR_ABM seeds it from
``pars.seed``, as the reference seeds numpy/numba via laser_core.random.seed,
model.py:82).
"""

import numpy as np

dist_list = ["constant", "exponential", "gamma", "lognormal", "normal", "poisson", "uniform"]
__all__ = [*dist_list, "Distribution", "lognormal_int", "set_seed", "get_rng"]

_rng = np.random.default_rng()


def set_seed(seed):
    global _rng
    _rng = np.random.default_rng(seed)
    return _rng


def get_rng():
    return _rng


class Distribution:
    _supported = {"constant", "exponential", "gamma", "lognormal", "lognormal_int", "normal", "poisson", "uniform"}

    def __init__(self, dist_type, **pars):
        if dist_type not in self._supported:
            raise ValueError(f"Unsupported distribution: {dist_type}. Supported: {sorted(self._supported)}")
        self.dist_type = dist_type
        self.pars = pars

    def sample(self, size=1):
        d, p, rng = self.dist_type, self.pars, _rng
        if d == "constant":
            return np.full(size, p.get("value", 1))
        if d == "exponential":
            return rng.exponential(p.get("scale", 1.0), size)
        if d == "gamma":
            return rng.gamma(p.get("shape", 2.0), p.get("scale", 1.0), size)
        if d in ("lognormal", "lognormal_int"):
            m, s = p.get("mean", 1.0), p.get("sigma", 0.5)
            if m > 0:
                mu = np.log(m**2 / np.sqrt(s**2 + m**2))
                sg = np.sqrt(np.log(s**2 / m**2 + 1))
                out = rng.lognormal(mean=mu, sigma=sg, size=size)
            else:
                out = np.zeros(size)
            return np.round(out) if d == "lognormal_int" else out
        if d == "normal":
            return rng.normal(p.get("mean", 0.0), p.get("std", 1.0), size)
        if d == "poisson":
            return rng.poisson(p.get("lam", 5), size)
        if d == "uniform":
            return rng.integers(p.get("min", 2), p.get("max", 10), size)
        raise ValueError(f"Unsupported distribution type: {d}")

    def mean(self):
        """Analytic mean of the distribution."""
        d, p = self.dist_type, self.pars
        if d == "constant":
            return float(p.get("value", 1))
        if d == "exponential":
            return float(p.get("scale", 1.0))
        if d == "gamma":
            return float(p.get("shape", 2.0) * p.get("scale", 1.0))
        if d in ("lognormal", "lognormal_int"):
            return float(p.get("mean", 1.0))
        if d == "normal":
            return float(p.get("mean", 0.0))
        if d == "poisson":
            return float(p.get("lam", 5))
        if d == "uniform":
            return (p.get("min", 2) + p.get("max", 10) - 1) / 2.0
        raise ValueError(d)

    def __call__(self, size=1):
        return self.sample(size)

    def __repr__(self):
        return f"Distribution(type={self.dist_type}, pars={self.pars})"


def constant(value):
    return Distribution("constant", value=value)


def exponential(scale):
    return Distribution("exponential", scale=scale)


def gamma(shape, scale):
    return Distribution("gamma", shape=shape, scale=scale)


def lognormal(mean, sigma):
    return Distribution("lognormal", mean=mean, sigma=sigma)


def lognormal_int(mean, sigma):
    return Distribution("lognormal_int", mean=mean, sigma=sigma)


def normal(mean, std):
    return Distribution("normal", mean=mean, std=std)


def poisson(lam):
    return Distribution("poisson", lam=lam)


def uniform(min, max):  # noqa: A002 -- reference signature
    return Distribution("uniform", min=min, max=max)
