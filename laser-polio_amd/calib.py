"""Calibration on the batched engine and on the mean-field adjoint.

The reference calibrates with Optuna (calib/calibrate.py, calib/objective.py:9-80): each
trial suggests parameters from a box (objective.py:18-28), runs ``n_replicates``
simulations one process at a time, reduces their results to summary targets and scores
them against observed data (calib/scoring.py:8-112), the trial's score being the mean
over replicates.  Optuna is not available here, and the GPU changes what a trial costs:

  * ``BatchedCalibration`` runs a whole generation of trials x replicates as one
    ``BatchedSims`` context (populations drawn on the GPU, every kernel launch of a day
    covering every simulation), scores each simulation with the reference's scoring
    functions, and moves the search by the cross-entropy method (elite fraction ->
    per-parameter normal, clipped to the box); gravity parameters can be applied to
    every simulation of a generation on the device (``BatchedSims.set_gravity``).
  * ``fit_meanfield`` fits the node map's parameters (r0_scalars, gravity k / a / b / c,
    seasonal amplitude and peak) by gradient descent on the mean-field loss, whose
    gradient the hand-written adjoint gives (adjoint.MeanFieldProblem).

``compute_fit`` and ``compute_log_likelihood_fit`` restate calib/scoring.py:8-112.
"""

import numpy as np


# -- scoring (calib/scoring.py) ----------------------------------------------------------------
def compute_fit(actual, predicted, use_squared=False, normalize=False, weights=None):
    """Weighted sum over targets of |actual - predicted| (squared, or normalised by the
    target's maximum, on request) (scoring.py:8-40)."""
    fit = 0.0
    weights = weights or {}
    for key in actual:
        if key not in predicted:
            continue
        v1 = np.asarray(actual[key], dtype=float)
        v2 = np.asarray(predicted[key], dtype=float)
        if v1.shape != v2.shape:
            continue
        g = np.abs(v1 - v2)
        if normalize and v1.max() > 0:
            g = g / v1.max()
        if use_squared:
            g = g**2
        fit += float((g * weights.get(key, 1)).sum())
    return fit


def compute_log_likelihood_fit(actual, predicted, method="poisson", dispersion=1.0, weights=None, norm_by_n=True):
    """Per-target negative log-likelihood of the observed values given the simulated ones
    (Poisson, or negative binomial with r = dispersion), simulated values clipped to 1e-6,
    weighted and divided by the number of values; nested dicts are flattened in sorted key
    order; plus "total_log_likelihood" = their sum (scoring.py:42-112)."""
    from scipy.stats import nbinom, poisson

    out = {}
    weights = weights or {}
    for key in actual:
        if key not in predicted:
            continue
        a, p = actual[key], predicted[key]
        if isinstance(a, dict) and isinstance(p, dict):
            if set(a) != set(p):
                continue
            ks = sorted(a)
            obs = np.array([a[k] for k in ks], dtype=float)
            sim = np.array([p[k] for k in ks], dtype=float)
        else:
            obs = np.asarray(a, dtype=float)
            sim = np.asarray(p, dtype=float)
        if obs.shape != sim.shape:
            continue
        sim = np.clip(sim, 1e-6, None)
        if method == "poisson":
            logp = poisson.logpmf(obs, sim)
        elif method == "neg_binomial":
            logp = nbinom.logpmf(obs, dispersion, dispersion / (dispersion + sim))
        else:
            raise ValueError(f"Unknown method '{method}'")
        logp = np.atleast_1d(logp)
        n = len(logp) if norm_by_n else 1
        out[key] = -1.0 * weights.get(key, 1.0) * float(logp.sum()) / n
    out["total_log_likelihood"] = sum(out.values())
    return out


def suggest_params(space, rng, n):
    """n parameter sets drawn uniformly from `space` = {name: {"low": .., "high": ..}}:
    integers when both bounds are ints, floats otherwise (objective.py:18-28)."""
    out = []
    for _ in range(n):
        d = {}
        for name, spec in space.items():
            lo, hi = spec["low"], spec["high"]
            if isinstance(lo, int) and isinstance(hi, int):
                d[name] = int(rng.integers(lo, hi + 1))
            elif isinstance(lo, float) or isinstance(hi, float):
                d[name] = float(rng.uniform(float(lo), float(hi)))
            else:
                raise TypeError(f"Cannot infer parameter type for '{name}'")
        out.append(d)
    return out


def default_targets(res, pars):
    """Summary targets of one simulation's results: total transmission infections, and
    infections per 30-day period and per node (results arrays [t, node])."""
    inc = np.asarray(res["new_exposed"], dtype=np.int64)[1:]
    for k in ("ri_protected", "sia_protected"):
        if k in res:
            inc = inc - np.asarray(res[k], dtype=np.int64)[1:]
    T = inc.shape[0] - inc.shape[0] % 30
    return {"total_infected": np.array([inc.sum()]),
            "monthly_infected": inc[:T].reshape(T // 30, 30, -1).sum(axis=(1, 2)),
            "node_infected": inc.sum(axis=0)}


# -- batched search -------------------------------------------------------------------------------
class BatchedCalibration:
    """Calibration whose trials run batched on the GPU.

    make_pars(params, replicate) -> PropertySet of one simulation (the model config merged
    with the suggested params, objective.py:37-42; the replicate picks its seed);
    target_fn(results, pars) -> summary targets (dict of arrays); actual: the observed
    targets; scoring_fn(actual, predicted, weights=...) -> dict with "total_log_likelihood";
    build(pars) -> SEIR_ABM (configs.build by default)."""

    def __init__(self, make_pars, space, actual, target_fn=default_targets, scoring_fn=compute_log_likelihood_fit,
                 n_replicates=1, weights=None, build=None, seed=0, gravity_on_device=False):
        self.make_pars, self.space, self.actual = make_pars, dict(space), actual
        self.target_fn, self.scoring_fn, self.weights = target_fn, scoring_fn, weights or {}
        self.n_replicates = int(n_replicates)
        self.rng = np.random.default_rng(seed)
        self.gravity_on_device = gravity_on_device
        self.trials = []
        if build is None:
            from . import configs

            build = configs.build
        self.build = build

    def evaluate(self, params_list):
        """Score every parameter set: n_replicates simulations each, all in one batch; a
        set's score is the mean of its replicates' total_log_likelihood (objective.py:45-80)."""
        from .batch import build_replicates_device

        plist, owner = [], []
        for i, prm in enumerate(params_list):
            for rep in range(self.n_replicates):
                plist.append(self.make_pars(prm, rep))
                owner.append(i)
        batch = build_replicates_device(plist, self.build)
        try:
            if self.gravity_on_device:
                ks = [p.gravity_k * 10 ** p.gravity_k_exponent for p in plist]
                batch.set_gravity(k=ks, a=[p.gravity_a for p in plist], b=[p.gravity_b for p in plist],
                                  c=[p.gravity_c for p in plist], max_rowsum=[p.max_migr_frac for p in plist],
                                  pull=False)
            eng = batch.prepare()
            eng.run_days(0, batch.cfg.nt, use_graph=True)
            eng.sync()
            per_set = [[] for _ in params_list]
            names = ("new_exposed", "ri_protected", "sia_protected")
            full = {}
            for attr in names:
                if attr in batch.arrays:
                    out = np.empty_like(batch.arrays[attr])
                    eng.pull(attr, out)
                    full[attr] = out
            for k, i in enumerate(owner):
                na, nb = batch.node_off[k], batch.node_off[k + 1]
                res = {a: v[:, na:nb] for a, v in full.items()}
                pred = self.target_fn(res, plist[k])
                score = self.scoring_fn(self.actual, pred, weights=self.weights)["total_log_likelihood"]
                per_set[i].append(float(score))
        finally:
            batch.close()
        out = []
        for prm, scores in zip(params_list, per_set):
            rec = {"params": dict(prm), "score": float(np.mean(scores)), "rep_scores": scores}
            self.trials.append(rec)
            out.append(rec)
        return out

    def run(self, n_generations=3, batch=32, elite_frac=0.25):
        """Cross-entropy search: generation 0 uniform over the box, then each generation drawn
        from per-parameter normals fitted to the previous generations' elite trials (lowest
        scores), clipped to the box.  Returns the best trial."""
        names = list(self.space)
        cand = suggest_params(self.space, self.rng, batch)
        for _ in range(n_generations):
            self.evaluate(cand)
            ranked = sorted(self.trials, key=lambda r: r["score"])
            elite = ranked[: max(2, int(round(elite_frac * len(ranked))))]
            cand = []
            for _ in range(batch):
                d = {}
                for nm in names:
                    lo, hi = self.space[nm]["low"], self.space[nm]["high"]
                    v = np.array([e["params"][nm] for e in elite], dtype=float)
                    sd = max(float(v.std()), 0.02 * (float(hi) - float(lo)))
                    x = float(np.clip(self.rng.normal(v.mean(), sd), lo, hi))
                    d[nm] = int(round(x)) if isinstance(lo, int) and isinstance(hi, int) else x
                cand.append(d)
        return min(self.trials, key=lambda r: r["score"])


# -- gradient fit on the mean-field map -----------------------------------------------------------
POSITIVE = ("r0s", "k", "a", "b", "c", "amp")


def fit_meanfield(problem, init=None, names=("r0s", "k", "amp"), steps=100, lr=0.05):
    """Fit `names` of an adjoint.MeanFieldProblem (target set) by Adam on the loss, positive
    parameters through their logarithm; the other parameters stay at `init` (default: the
    simulation's own values).  Every step is one forward and one hand-written adjoint pass.
    Returns (fitted params, loss per step)."""
    import torch

    p0 = problem.params() if init is None else {k: v.detach().clone() for k, v in init.items()}
    theta = {}
    for k in names:
        v = p0[k]
        theta[k] = (torch.log(v) if k in POSITIVE else v.clone()).requires_grad_(True)
    opt = torch.optim.Adam(list(theta.values()), lr=lr)
    history = []

    def current():
        q = dict(p0)
        for k, t in theta.items():
            q[k] = torch.exp(t) if k in POSITIVE else t
        return q

    for _ in range(steps):
        opt.zero_grad()
        L = problem.loss(current())
        L.backward()
        opt.step()
        history.append(float(L))
    with torch.no_grad():
        final = {k: v.detach() for k, v in current().items()}
    return final, history
