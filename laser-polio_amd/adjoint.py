"""Differentiable mean-field rollout of the node-level transmission map.

The stochastic ABM stays the model of record; this module gives gradients of a
calibration loss with respect to the parameters of its node-level force of
infection (Transmission_ABM.step Part 2b, src/laser_polio/model.py:1327-1363):
per-node r0_scalars, the gravity network parameters k, a, b, c
(model.py:1243-1258) and the seasonality amplitude and peak day
(utils.py:589-625).

Node compartments are aggregated from an ABM population at day 0 and rolled
forward deterministically (expected exposures, daily E -> I and I -> R exits at
the mean durations).  ``MeanFieldRollout`` is a torch.autograd.Function whose
forward and hand-written adjoint run in liblpstep.so (lp_meanfield.h); it keeps
the state only every ``checkpoint`` days and recomputes each segment forward
before its adjoint.  ``rollout_reference`` is the same rollout in plain PyTorch
ops, for autograd and finite-difference checks.
"""

import ctypes

import numpy as np
import torch

from . import engine as engmod


class LPMFModel(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("net", ctypes.c_void_p),
        ("net_t", ctypes.c_void_p),
        ("keep", ctypes.c_void_p),
        ("r0_scalar", ctypes.c_void_p),
        ("rho", ctypes.c_void_p),
        ("inv_pop", ctypes.c_void_p),
        ("season", ctypes.c_void_p),
        ("iota", ctypes.c_float),
        ("inv_tau_e", ctypes.c_float),
        ("inv_tau_i", ctypes.c_float),
        ("pad", ctypes.c_float),
    ]


_mf = None


def _lib():
    global _mf
    if _mf is None:
        lib = engmod.load_library()
        vp, i32 = ctypes.c_void_p, ctypes.c_int32
        lib.lp_mf_days.argtypes = [ctypes.POINTER(LPMFModel), i32, i32] + [vp] * 7
        lib.lp_mf_adjoint_days.argtypes = [ctypes.POINTER(LPMFModel), i32, i32] + [vp] * 13
        _mf = lib
    return _mf


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _check(rc, what):
    if rc != 0:
        raise engmod.LPError(f"{what} failed ({engmod.ERRORS.get(rc, rc)})")


class MeanFieldRollout(torch.autograd.Function):
    """inc[t, i] = expected new exposures of node i on day t, t < len(season)."""

    @staticmethod
    def forward(ctx, r0s, net, season, S0, E0, I0, rho, inv_pop, iota, inv_tau_e, inv_tau_i, checkpoint=30):
        dev = r0s.device
        if dev.type != "cuda":
            raise engmod.EngineUnavailable("the mean-field rollout runs on the HIP path (GPU tensors)")
        f32 = dict(dtype=torch.float32, device=dev)
        r0s, season = r0s.detach().to(**f32).contiguous(), season.detach().to(**f32).contiguous()
        net = net.detach().to(**f32).contiguous()
        net_t = net.t().contiguous()
        keep = (1.0 - net.sum(dim=1)).contiguous()
        rho, inv_pop = rho.to(**f32).contiguous(), inv_pop.to(**f32).contiguous()
        n, T = r0s.numel(), season.numel()
        model = LPMFModel(n=n, net=net.data_ptr(), net_t=net_t.data_ptr(), keep=keep.data_ptr(),
                          r0_scalar=r0s.data_ptr(), rho=rho.data_ptr(), inv_pop=inv_pop.data_ptr(),
                          season=season.data_ptr(), iota=float(iota), inv_tau_e=float(inv_tau_e),
                          inv_tau_i=float(inv_tau_i))
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        inc = torch.empty((T, n), **f32)
        state = [x.detach().to(**f32).reshape(1, n).contiguous() for x in (S0, E0, I0)]
        checkpoints = []
        for t0 in range(0, T, checkpoint):
            t1 = min(T, t0 + checkpoint)
            checkpoints.append((t0, t1, [x.clone() for x in state]))
            traj = [torch.empty((t1 - t0 + 1, n), **f32) for _ in range(3)]
            for tr, x in zip(traj, state):
                tr[0].copy_(x[0])
            _check(_lib().lp_mf_days(ctypes.byref(model), t0, t1, _p(traj[0]), _p(traj[1]), _p(traj[2]),
                                     _p(inc[t0:t1]), None, None, stream), "lp_mf_days")
            state = [tr[-1:].clone() for tr in traj]
        ctx.model = model
        ctx.keepalive = (net, net_t, keep, r0s, rho, inv_pop, season)
        ctx.checkpoints = checkpoints
        ctx.n, ctx.T = n, T
        return inc

    @staticmethod
    def backward(ctx, g_inc):
        net, net_t, keep, r0s, rho, inv_pop, season = ctx.keepalive
        model, n, T = ctx.model, ctx.n, ctx.T
        dev = r0s.device
        f32 = dict(dtype=torch.float32, device=dev)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        g_inc = g_inc.to(**f32).contiguous()
        r0_bar = torch.zeros(n, **f32)
        net_bar = torch.zeros((n, n), **f32)
        season_part = torch.zeros((T, n), **f32)
        aS, aE, aI = (torch.zeros(n, **f32) for _ in range(3))
        work = torch.zeros(2 * n, **f32)
        for t0, t1, state in reversed(ctx.checkpoints):
            L = t1 - t0
            traj = [torch.empty((L + 1, n), **f32) for _ in range(3)]
            for tr, x in zip(traj, state):
                tr[0].copy_(x[0])
            inc = torch.empty((L, n), **f32)
            msave, gsave = torch.empty((L, n), **f32), torch.empty((L, n), **f32)
            _check(_lib().lp_mf_days(ctypes.byref(model), t0, t1, _p(traj[0]), _p(traj[1]), _p(traj[2]), _p(inc),
                                     _p(msave), _p(gsave), stream), "lp_mf_days")
            _check(_lib().lp_mf_adjoint_days(ctypes.byref(model), t0, t1, _p(traj[0]), _p(traj[2]), _p(msave),
                                             _p(gsave), _p(g_inc[t0:t1]), _p(aS), _p(aE), _p(aI), _p(r0_bar),
                                             _p(net_bar), _p(season_part[t0:t1]), _p(work), stream),
                   "lp_mf_adjoint_days")
        season_bar = season_part.sum(dim=1)
        return r0_bar, net_bar, season_bar, aS, aE, aI, None, None, None, None, None, None


def rollout(r0s, net, season, S0, E0, I0, rho, inv_pop, iota, inv_tau_e, inv_tau_i, checkpoint=30):
    return MeanFieldRollout.apply(r0s, net, season, S0, E0, I0, rho, inv_pop, iota, inv_tau_e, inv_tau_i, checkpoint)


def rollout_reference(r0s, net, season, S0, E0, I0, rho, inv_pop, iota, inv_tau_e, inv_tau_i):
    """The same rollout in plain PyTorch ops (any dtype / device), for autograd and FD checks."""
    keep = 1.0 - net.sum(dim=1)
    S, E, I = S0, E0, I0
    out = []
    for t in range(season.shape[0]):
        m = iota * (I * keep + net.t() @ I)
        g = m * season[t] * r0s * inv_pop
        lam = rho * S * (-torch.expm1(-g))
        S, E, I = S - lam, E + lam - E * inv_tau_e, I + E * inv_tau_e - I * inv_tau_i
        out.append(lam)
    return torch.stack(out)


# -- parameterisation (differentiable, plain torch) ------------------------------------------
def gravity_network(pop, dist, k, a, b, c, max_frac, k_exponent=0.0):
    """Row-normalised gravity network as Transmission_ABM builds it (model.py:1243-1258):
    k 10^kexp p_i^a p_j^b / d_ij^c, zero diagonal, divided by (sum p)^c, rows with a sum
    above max_frac scaled down to max_frac."""
    n = pop.shape[0]
    off = ~torch.eye(n, dtype=torch.bool, device=pop.device)
    d = torch.where(off, dist, torch.ones_like(dist))
    net = k * (10.0 ** k_exponent) * pop[:, None] ** a * pop[None, :] ** b / d ** c
    net = torch.where(off, net, torch.zeros_like(net)) / pop.sum() ** c
    rs = net.sum(dim=1, keepdim=True)
    scale = torch.where(rs > max_frac, max_frac / rs, torch.ones_like(rs))
    return net * scale


def seasonality(amp, peak_doy, doy, days_in_year):
    """1 + amp cos(2 pi (doy - peak) / days_in_year) (utils.py:589-625), doy 1-based."""
    return 1.0 + amp * torch.cos(2.0 * np.pi * (doy - peak_doy) / days_in_year)


def weekly(inc, week=7):
    T = inc.shape[0] - inc.shape[0] % week
    return inc[:T].reshape(T // week, week, -1).sum(dim=1)


def poisson_nll(mu, y, eps=1e-12):
    """Negative Poisson log-likelihood (without the log y! term) of counts y given means mu."""
    return (mu - y * torch.log(mu + eps)).sum()


def aggregate(sim):
    """Node compartments and constants of a constructed SEIR_ABM at day 0: S, E, I counts,
    mean acquisition risk of the susceptibles, 1/max(pop,1), mean infectious weight per
    infectious agent (daily_infectivity x strain-0 r0 scalar) and mean E / I durations."""
    p, pars = sim.people, sim.pars
    N = len(sim.nodes)
    st = p.disease_state
    node = p.node_id
    live = st >= 0
    S = np.bincount(node[live & (st == 0)], minlength=N).astype(np.float64)
    E = np.bincount(node[live & (st == 1)], minlength=N).astype(np.float64)
    I = np.bincount(node[live & (st == 2)], minlength=N).astype(np.float64)
    risk = p.acq_risk_multiplier
    sw = np.bincount(node[live & (st == 0)], weights=risk[live & (st == 0)], minlength=N)
    rho = np.where(S > 0, sw / np.maximum(S, 1), 1.0)
    pop = np.bincount(node[live], minlength=N).astype(np.float64)
    inf = p.daily_infectivity[live]
    iota = float(inf.mean()) * float(list(pars.strain_r0_scalars.values())[0])
    return {
        "S": S, "E": E, "I": I, "rho": rho, "inv_pop": 1.0 / np.maximum(pop, 1.0), "pop": pop,
        "iota": iota, "inv_tau_e": 1.0 / max(1e-6, float(pars.dur_exp.mean())),
        "inv_tau_i": 1.0 / max(1e-6, float(pars.dur_inf.mean())),
    }


class MeanFieldProblem:
    """The mean-field calibration loss of a constructed SEIR_ABM: node compartments
    aggregated from its agents when the problem is made (aggregate: make it before the
    simulation runs), its geography, seasonality calendar and gravity settings, T days;
    with a target set (set_target: daily per-node incidence [T, nodes]), loss(params) is
    the Poisson negative log-likelihood of weekly incidence of the T-day rollout.
    params: r0s (per node), k, a, b, c (gravity), amp, peak (seasonality)."""

    NAMES = ("r0s", "k", "a", "b", "c", "amp", "peak")

    def __init__(self, sim, T, dtype=torch.float32, device="cuda"):
        import calendar

        agg = aggregate(sim)
        rp = sim.pars
        tx = sim._instance("Transmission_ABM")
        self.agg, self.rp, self.T = agg, rp, int(T)
        dist = getattr(tx, "distances", None)
        dist = np.asarray(rp.distances if dist is None else dist, dtype=np.float64)
        doy = np.array([sim.datevec[t].timetuple().tm_yday for t in range(1, T + 1)], dtype=np.float64)
        diy = np.array([366.0 if calendar.isleap(sim.datevec[t].year) else 365.0 for t in range(1, T + 1)])
        self._np = {"doy": doy, "diy": diy, "pop": np.asarray(rp.init_pop, dtype=np.float64), "dist": dist, "y": None}
        self._base = {"r0s": np.asarray(tx.r0_scalars, dtype=np.float64) * np.ones(len(rp.init_pop)),
                      "k": float(rp.gravity_k), "a": float(rp.gravity_a), "b": float(rp.gravity_b),
                      "c": float(rp.gravity_c), "amp": float(rp.seasonal_amplitude), "peak": float(rp.seasonal_peak_doy)}
        self.y = None
        self.to(dtype, device)

    def _t(self, x):
        return torch.as_tensor(np.asarray(x, dtype=np.float64), dtype=self.dtype, device=self.device)

    def to(self, dtype, device):
        self.dtype, self.device = dtype, device
        self.const = {k: self._t(self.agg[k]) for k in ("S", "E", "I", "rho", "inv_pop")}
        self.doy, self.diy, self.pop, self.dist = (self._t(self._np[k]) for k in ("doy", "diy", "pop", "dist"))
        self.y = None if self._np["y"] is None else weekly(self._t(self._np["y"]))
        return self

    def set_target(self, y_daily):
        y = np.asarray(y_daily, dtype=np.float64)
        if y.shape != (self.T, len(self._np["pop"])):
            raise ValueError(f"target must be [T={self.T}, nodes], got {y.shape}")
        self._np["y"] = y
        self.y = weekly(self._t(y))
        return self

    def params(self, **scale):
        """The simulation's own parameter values as tensors, each multiplied by scale[name]."""
        return {k: self._t(np.asarray(v, dtype=np.float64) * scale.get(k, 1.0)) for k, v in self._base.items()}

    def incidence(self, params, fn=rollout):
        rp, agg = self.rp, self.agg
        net = gravity_network(self.pop, self.dist, params["k"], params["a"], params["b"], params["c"],
                              float(rp.max_migr_frac), float(rp.gravity_k_exponent))
        season = seasonality(params["amp"], params["peak"], self.doy, self.diy)
        return fn(params["r0s"], net, season, self.const["S"], self.const["E"], self.const["I"], self.const["rho"],
                  self.const["inv_pop"], agg["iota"], agg["inv_tau_e"], agg["inv_tau_i"])

    def loss(self, params, fn=rollout):
        if self.y is None:
            raise ValueError("set_target before loss")
        return poisson_nll(weekly(self.incidence(params, fn)), self.y)
