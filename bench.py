"""Benchmark: agent-days/s of the laser-polio daily step on one MI355X.

Workload (the national-scale configuration): 774 nodes, populations
LogNormal(ln 2.5e5, 0.6) -> about 2.3e8 agents, gravity network, RI + SIA
campaigns, 2555 simulated days.  A "step" is one simulated day of the whole
population.  The host initialises the population (not timed), uploads it once,
runs ``--warmup`` days untimed, then times exactly ``--steps`` days bracketed by
torch.cuda.synchronize().  value = sum over timed days of live agents / seconds.

Also reported on the same JSON line:
  roofline      the agent pass (k_ei every day, plus k_full / k_ri_list on VD / RI / SIA days):
                algorithmic bytes per day (agent_pass_bytes below) over its
                HIP-event-timed duration, against MI355X's 8 TB/s HBM peak
  cpu_baseline  the CPU oracle (restated reference kernels, OpenMP on all host
                threads) on the same initial population for the first --cpu-days days,
                plus one further day on a single thread
The HBM traffic of the agent pass (rocprofv3 FETCH_SIZE + WRITE_SIZE per launch,
calibrated on a plain copy) is read from profiles/agent_pass_national_traffic.md
when that file was measured on the same workload and window; otherwise null.
"""

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "agent-days/s, 200M-agent Nigeria run; replicate-sims/hr batched vs serial"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def agent_pass_bytes(res, flags, t0, t1):
    """Algorithmic bytes of the agent pass over ticks [t0, t1).

    DiseaseState and the E / I tally cost 17 B per E or I agent: its timers read and
    written (2 x (2 + 2)), strain (1), infectivity (4) and the paralysis fields (4); the
    state bytes of S and R agents are not compulsory once E and I agents are found
    without a scan.  VD step days add the per-chunk earliest death dates (4 B per 256
    slots), 6 B per death (date_of_death, potentially_paralyzed, paralyzed) and 17 B per
    birth (date_of_birth, date_of_death, node_id, ri_timer and state written, risk read).
    RI and SIA days scan every agent's state and chronically_missed (2 B); RI days add
    ri_timer (2 B, read: the decrement is applied by arithmetic), SIA days date_of_birth
    (4 B).  N_live = live agents that day (results.pop); E and I are the counts at the
    start of the day (the previous row).
    """
    from laser_polio_amd import engine as em

    pop = res["pop"].astype(np.int64).sum(axis=1)
    n_e = res["E"].astype(np.int64).sum(axis=1)
    n_i = res["I"].astype(np.int64).sum(axis=1)
    births = res["births"].astype(np.int64).sum(axis=1) if "births" in res else np.zeros_like(pop)
    deaths = res["deaths"].astype(np.int64).sum(axis=1) if "deaths" in res else np.zeros_like(pop)
    total = 0
    for t in range(t0, t1):
        n_live, f = int(pop[t]), int(flags[t])
        b = 17 * (int(n_e[t - 1]) + int(n_i[t - 1]))
        if f & em.DAY_VD_STEP:
            b += 4 * (n_live // 256) + 6 * int(deaths[t]) + 17 * int(births[t])
        if f & (em.DAY_RI | em.DAY_SIA):
            b += 2 * n_live
        if f & em.DAY_RI:
            b += 2 * n_live
        if f & em.DAY_SIA:
            b += 4 * n_live
        total += b
    return total


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "agent_pass_national_traffic.md")


def workload_key(scale, warmup, steps, tpw):
    return f"scale={scale:g},warmup={warmup},steps={steps},tiles_per_block={tpw}"


def measured_traffic(key, path=TRAFFIC_PROFILE):
    """HBM bytes per agent-pass launch from the committed rocprofv3 summary, for this workload only."""
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic: (\{.*\})$", open(path).read(), re.M)
    if not m:
        return None
    rec = json.loads(m.group(1))
    return rec if rec.get("workload") == key else None


def dump_days(path, res, flags, t0, t1):
    """Per-day flags, N_live, E, I and strain-0 I (start of day), N_draw (agents of nodes
    with transmission draws), births and deaths."""
    pop = res["pop"].astype(np.int64)
    n_e, n_i = res["E"].astype(np.int64).sum(axis=1), res["I"].astype(np.int64).sum(axis=1)
    n_i0 = res["I_by_strain"][:, :, 0].astype(np.int64).sum(axis=1)
    tx_new = res["new_exposed"].astype(np.int64) - res["ri_protected"] - res["sia_protected"]
    with open(path, "w") as f:
        f.write("t,flags,n_live,n_e,n_i,n_i0,n_draw,births,deaths\n")
        for t in range(t0, t1):
            n_draw = int(pop[t][tx_new[t] > 0].sum())
            f.write(f"{t},{int(flags[t])},{int(pop[t].sum())},{int(n_e[t - 1])},{int(n_i[t - 1])},"
                    f"{int(n_i0[t - 1])},{n_draw},"
                    f"{int(res['births'][t].sum())},{int(res['deaths'][t].sum())}\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2550)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=float, default=1.0, help="population scale (1.0 = ~2.3e8 agents)")
    ap.add_argument("--cpu-days", type=int, default=24, help="days of the CPU oracle baseline sample (0 = skip)")
    ap.add_argument("--dump-days", default="", help="write per-day N_live / N_EI / N_draw / flags to this CSV")
    ap.add_argument("--graph", type=int, default=0, help="1: replay one hipGraph per day (no per-kernel events)")
    ap.add_argument("--tiles-per-block", type=int, default=0)
    ap.add_argument("--workload", choices=["national", "replicates", "adjoint"], default="national",
                    help="national: one 774-node simulation (default); replicates: --sims independent "
                         "simulations of uneven sizes batched in one context, against serial runs of a sample; "
                         "adjoint: gradients of a mean-field calibration loss on the national geography")
    ap.add_argument("--sims", type=int, default=512, help="replicates: simulations in the batch")
    ap.add_argument("--serial-sample", type=int, default=16, help="replicates: simulations also run one at a time")
    ap.add_argument("--workers", type=int, default=16, help="replicates: host processes initialising simulations")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="instead of the benchmark: copy this many GiB device-to-device 5 times (a plain copy kernel "
                         "of known size, to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        import torch

        n = int(args.copy_calibration * (1 << 30))
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "copies": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding the nodes over GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    if args.workload == "replicates":
        replicates(args)
        return
    if args.workload == "adjoint":
        adjoint_workload(args)
        return

    import torch

    lp = __graft_entry__.load_package()
    from laser_polio_amd import configs

    nt_needed = 1 + args.warmup + args.steps
    t_init = time.perf_counter()
    pars = configs.nigeria_national_pars(scale=args.scale, dur=max(2555, nt_needed - 1))
    if args.tiles_per_block:
        pars.tiles_per_block = args.tiles_per_block
    sim = configs.build(pars)
    n0 = int(pars.init_pop.sum())
    log(f"[bench] init: {len(pars.init_pop)} nodes, {n0} agents, capacity {sim.people.capacity}, "
        f"{time.perf_counter() - t_init:.1f}s")

    oracle_sim = None
    if args.cpu_days > 0:
        from oracle.oracle import OracleSim

        oracle_sim = OracleSim(sim.export_tables(), pop0=sim.pars.init_pop)

    sim._timing = not args.graph
    t_up = time.perf_counter()
    eng = sim._ensure_engine()
    torch.cuda.synchronize()
    log(f"[bench] upload {time.perf_counter() - t_up:.1f}s")
    day_flags = eng.arrays["day_flags"]

    eng.run_days(0, 1 + args.warmup, use_graph=bool(args.graph))
    eng.sync()
    eng.agent_pass_timing()  # discard warmup events
    torch.cuda.synchronize()
    t_begin, t_end = 1 + args.warmup, 1 + args.warmup + args.steps
    t0 = time.perf_counter()
    eng.run_days(t_begin, t_end, use_graph=bool(args.graph))
    eng.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    launches, kernel_ms = (0, 0.0) if args.graph else eng.agent_pass_timing()
    sim._mark_device_newer()
    res = {k: getattr(sim.results, k) for k in ("pop", "E", "I", "I_by_strain", "new_exposed", "ri_protected",
                                                 "sia_protected", "births", "deaths")}
    agent_days = int(res["pop"][t_begin:t_end].astype(np.int64).sum())
    value = agent_days / elapsed
    log(f"[bench] {args.steps} days in {elapsed:.3f}s -> {value:.4e} agent-days/s; "
        f"exposures {int(res['new_exposed'][t_begin:t_end].sum())}")

    if args.dump_days:
        dump_days(args.dump_days, res, day_flags, t_begin, t_end)
    roofline = None
    if launches:
        bytes_alg = agent_pass_bytes(res, day_flags, t_begin, t_end)
        achieved = bytes_alg / (kernel_ms / 1e3) / 1e9
        tpw = int(sim.pars.tiles_per_block or 16)
        tr = measured_traffic(workload_key(args.scale, args.warmup, args.steps, tpw))
        roofline = {
            "bound": "hbm",
            "kernel": "agent pass (k_ei + k_full / k_ri_list)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (tr["bytes_per_launch"] if tr else None),
            "launches": launches,
            "avg_launch_us": round(kernel_ms * 1e3 / launches, 2),
            "bytes_per_launch": int(bytes_alg // launches),
        }
        if tr:
            roofline["traffic_over_algorithmic"] = round(tr["bytes_per_launch"] / (bytes_alg / launches), 3)
        log(f"[bench] agent pass: {launches} launches, avg {kernel_ms * 1e3 / launches:.1f} us, {achieved:.0f} GB/s")

    cpu = None
    if oracle_sim is not None:
        from oracle import oracle as orc

        days = min(args.cpu_days, sim.nt - 1)
        oracle_sim.day(0)
        tc = time.perf_counter()
        for t in range(1, 1 + days):
            oracle_sim.day(t)
        cpu_s = time.perf_counter() - tc
        cpu_ad = int(oracle_sim.res["pop"][1 : 1 + days].astype(np.int64).sum())
        threads = int(orc.lib().orc_max_threads())
        single = None
        if days + 1 < sim.nt:
            oracle_sim.m.n_threads = 1
            tc = time.perf_counter()
            oracle_sim.day(days + 1)
            one_s = time.perf_counter() - tc
            single = {"value": round(int(oracle_sim.res["pop"][days + 1].astype(np.int64).sum()) / one_s, 1), "cores": 1,
                      "sample": f"day {days + 1} ({one_s:.1f}s)"}
        cpu = {
            "value": round(cpu_ad / cpu_s, 1),
            "unit": "agent-days/s",
            "cores": threads,
            "kind": "port",
            "sample": f"days 1-{days} of the same {n0}-agent population ({cpu_s:.1f}s, OpenMP {threads} threads, "
                      f"host nproc {os.cpu_count()})",
            "single_thread": single,
        }
        log(f"[bench] cpu oracle: {cpu['value']:.4e} agent-days/s over {days} days ({cpu_s:.1f}s)")

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "agent-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"national: 774 nodes, {n0} agents (population scale {args.scale}), RI + SIA every 180 d, "
                        f"days {t_begin}-{t_end - 1} of {sim.nt - 1} timed",
            "nodes": int(len(pars.init_pop)),
            "agents": n0,
            "capacity": int(sim.people.capacity),
            "graph": bool(args.graph),
            "arithmetic": "int8 state, uint16 timers, f32 weights -> int64 fixed-point sums, f32 node map",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def replicates(args):
    """Batched replicates: `--sims` simulations of the replicate workload (node count 8-40,
    0.5-6 M agents each at scale 1, R0 8-18, seasonality amplitude 0-0.6, gravity k 0.1-1,
    c 1.5-2.5, distinct seeds, 365 days) in one library context, every kernel launch of a
    day covering all of them; against the same simulations run one at a time (one context
    each, the same kernels and hipGraphs).  Serial time is measured on an evenly spaced
    sample and extended to the whole set by a least-squares fit of time = a * days + b *
    agent-days over the sample.  Host initialisation and uploads are not timed."""
    lp = __graft_entry__.load_package()
    from laser_polio_amd import configs
    from laser_polio_amd.batch import build_replicates

    S, dur = args.sims, 365
    pars_list = [configs.replicate_pars(i, scale=args.scale, dur=dur) for i in range(S)]
    t_init = time.perf_counter()
    batch = build_replicates(pars_list, configs.build, workers=max(1, args.workers), log=log)
    n0 = int(sum(int(p.init_pop.sum()) for p in pars_list))
    log(f"[bench] replicates: {S} simulations, {batch.cfg.n_nodes} nodes, {n0} agents, capacity "
        f"{batch.cfg.capacity}, host init {time.perf_counter() - t_init:.1f}s")

    import torch

    try:
        if args.tiles_per_block:
            tpw, tune = args.tiles_per_block, None
            batch.tiles_per_block = tpw
            batch.cfg.tiles_per_block = tpw
        else:
            tpw, tune = batch.autotune()
            log(f"[bench] autotune tiles_per_block: " + ", ".join(f"{k}: {v * 1e3:.1f} ms" for k, v in tune.items())
                + f" -> {tpw}")
        sample = sorted(set(np.linspace(0, S - 1, min(args.serial_sample, S)).round().astype(int).tolist()))
        singles = [batch.subset([k], tiles_per_block=tpw) for k in sample]
        t_up = time.perf_counter()
        eng = batch.prepare()
        torch.cuda.synchronize()
        log(f"[bench] upload {time.perf_counter() - t_up:.1f}s")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.run_days(0, batch.cfg.nt, use_graph=True)
        eng.sync()
        torch.cuda.synchronize()
        el_b = time.perf_counter() - t0
        ad_b = batch.agent_days(1, batch.cfg.nt)
        per_sim_ad = []
        pop = np.empty_like(batch.arrays["pop"])
        eng.pull("pop", pop)
        for k in range(S):
            per_sim_ad.append(int(pop[1:, batch.node_off[k]:batch.node_off[k + 1]].astype(np.int64).sum()))
        batch.close()
        log(f"[bench] batched: {S} sims x {dur} days in {el_b:.3f}s -> {S / el_b * 3600:.4e} sims/hr, "
            f"{ad_b / el_b:.4e} agent-days/s")

        times, ads = [], []
        for k, one in zip(sample, singles):
            e1 = one.prepare()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e1.run_days(0, one.cfg.nt, use_graph=True)
            e1.sync()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            ads.append(one.agent_days(1, one.cfg.nt))
            one.close()
        A = np.vstack([np.full(len(times), dur, float), np.array(ads, float)]).T
        coef, *_ = np.linalg.lstsq(A, np.array(times), rcond=None)
        serial_total = float(coef[0] * dur * S + coef[1] * sum(per_sim_ad))
        serial_total = max(serial_total, float(np.sum(times)))
        log(f"[bench] serial sample: {len(times)} sims in {np.sum(times):.3f}s ({np.mean(times) * 1e3:.1f} ms each); "
            f"fit {coef[0] * 1e6:.1f} us/day + {coef[1] * 1e12:.2f} ps/agent-day -> {serial_total:.2f}s for all {S}")
    finally:
        batch.close()
    out = {
        "metric": "replicate-sims/hr batched vs serial (replicate workload, 365 days)",
        "value": round(S / el_b * 3600, 1),
        "unit": "sims/hr",
        "n_gpus": 1,
        "steps": dur,
        "warmup": 0,
        "ms_per_step": round(el_b * 1e3 / dur, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"replicates: {S} simulations, {batch.cfg.n_nodes} nodes, {n0} agents "
                               f"(scale {args.scale}), 365 days, batched in one context",
                   "tiles_per_block": int(tpw)},
        "serial_sims_per_hr": round(S / serial_total * 3600, 1),
        "speedup_vs_serial": round(serial_total / el_b, 2),
        "batched_agent_days_per_s": round(ad_b / el_b, 1),
        "serial_sample": {"sims": len(times), "seconds": round(float(np.sum(times)), 4),
                          "fit_us_per_day": round(float(coef[0]) * 1e6, 2),
                          "fit_ps_per_agent_day": round(float(coef[1]) * 1e12, 3)},
    }
    print(json.dumps(out), flush=True)


def adjoint_workload(args):
    """Gradients of the mean-field calibration loss on the national geography at `--scale`
    (0.1 when left at 1.0: 774 nodes, about 2.3e7 agents): node compartments aggregated from
    the agent population at day 0, 180 days rolled forward, Poisson negative log-likelihood
    of weekly per-node incidence against the weekly transmission incidence of the stochastic
    ABM run on the GPU at the known parameters; gradient at perturbed parameters (r0_scalars
    x 0.9, gravity k x 0.8, amplitude x 1.2) with respect to the 774 r0_scalars, gravity
    k, a, b, c and seasonal amplitude and peak.  Reported: seconds per loss + gradient
    (HIP forward and adjoint, 30-day checkpoints), relative error against autograd of the
    plain float32 PyTorch rollout (per parameter group) and against float64 central
    differences (h = 1e-4 relative) on scalars and a sample of r0_scalars."""
    import torch

    __graft_entry__.load_package()
    from laser_polio_amd import adjoint as adj
    from laser_polio_amd import configs
    import calendar

    scale = 0.1 if args.scale == 1.0 else args.scale
    T = 180
    pars = configs.nigeria_national_pars(scale=scale, dur=T)
    sim = configs.build(pars)
    agg = adj.aggregate(sim)
    tx = sim._instance("Transmission_ABM")
    rp = sim.pars
    doy = np.array([sim.datevec[t].timetuple().tm_yday for t in range(1, T + 1)], dtype=np.float64)
    diy = np.array([366.0 if calendar.isleap(sim.datevec[t].year) else 365.0 for t in range(1, T + 1)])
    dist = np.asarray(rp.distances, dtype=np.float64)
    log(f"[bench] adjoint: {len(rp.init_pop)} nodes, {int(rp.init_pop.sum())} agents (scale {scale}), {T} days")
    t0 = time.perf_counter()
    sim.run()
    y_daily = (sim.results.new_exposed.astype(np.int64) - sim.results.ri_protected - sim.results.sia_protected)[1:T + 1]
    log(f"[bench] ABM target: {int(y_daily.sum())} transmission exposures in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda", 0)

    def tensors(dtype, device):
        t = lambda x: torch.as_tensor(np.asarray(x, dtype=np.float64), dtype=dtype, device=device)  # noqa: E731
        const = {k: t(agg[k]) for k in ("S", "E", "I", "rho", "inv_pop")}
        params = {"r0s": t(np.asarray(tx.r0_scalars, dtype=np.float64) * 0.9), "k": t(rp.gravity_k * 0.8),
                  "a": t(rp.gravity_a), "b": t(rp.gravity_b), "c": t(rp.gravity_c),
                  "amp": t(rp.seasonal_amplitude * 1.2), "peak": t(float(rp.seasonal_peak_doy))}
        return const, params, t(doy), t(diy), t(rp.init_pop), t(dist)

    def loss(params, const, doy_t, diy_t, pop_t, dist_t, y, fn):
        net = adj.gravity_network(pop_t, dist_t, params["k"], params["a"], params["b"], params["c"],
                                  float(rp.max_migr_frac), float(rp.gravity_k_exponent))
        season = adj.seasonality(params["amp"], params["peak"], doy_t, diy_t)
        inc = fn(params["r0s"], net, season, const["S"], const["E"], const["I"], const["rho"], const["inv_pop"],
                 agg["iota"], agg["inv_tau_e"], agg["inv_tau_i"])
        return adj.poisson_nll(adj.weekly(inc), y)

    c32, p32, doy32, diy32, pop32, dist32 = tensors(torch.float32, dev)
    y32 = adj.weekly(torch.as_tensor(y_daily, dtype=torch.float32, device=dev))

    def grad_hip():
        p = {k: v.clone().requires_grad_(True) for k, v in p32.items()}
        L = loss(p, c32, doy32, diy32, pop32, dist32, y32, adj.rollout)
        L.backward()
        return float(L), {k: v.grad.detach() for k, v in p.items()}

    grad_hip()
    torch.cuda.synchronize()
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        L_hip, g_hip = grad_hip()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    p = {k: v.clone().requires_grad_(True) for k, v in p32.items()}
    t0 = time.perf_counter()
    L_ag = loss(p, c32, doy32, diy32, pop32, dist32, y32, adj.rollout_reference)
    L_ag.backward()
    torch.cuda.synchronize()
    t_ag = time.perf_counter() - t0
    rel_ag = {k: float((g_hip[k].double() - p[k].grad.double()).norm() / p[k].grad.double().norm().clamp_min(1e-30))
              for k in p32}
    c64, p64, doy64, diy64, pop64, dist64 = tensors(torch.float64, "cpu")
    y64 = y32.double().cpu()
    f64 = lambda q: float(loss(q, c64, doy64, diy64, pop64, dist64, y64, adj.rollout_reference))  # noqa: E731
    fd_err = {}
    for name, idx in [("k", 0), ("a", 0), ("b", 0), ("c", 0), ("amp", 0), ("peak", 0)] + [("r0s", i) for i in (0, 100, 400)]:
        base = {k: v.clone() for k, v in p64.items()}
        x0 = float(base[name].reshape(-1)[idx])
        h = 1e-4 * max(abs(x0), 1.0)
        plus = {k: v.clone() for k, v in base.items()}
        minus = {k: v.clone() for k, v in base.items()}
        plus[name].reshape(-1)[idx] += h
        minus[name].reshape(-1)[idx] -= h
        fd = (f64(plus) - f64(minus)) / (2 * h)
        g = float(g_hip[name].reshape(-1)[idx])
        fd_err[f"{name}[{idx}]" if name == "r0s" else name] = abs(g - fd) / max(abs(fd), 1e-12)
    ms = float(np.median(times)) * 1e3
    log(f"[bench] loss {L_hip:.6e}; HIP forward+adjoint {ms:.2f} ms (autograd of plain float32 rollout {t_ag * 1e3:.1f} ms)")
    log(f"[bench] relative error vs autograd: {max(rel_ag.values()):.2e}; vs central differences: {max(fd_err.values()):.2e}")
    out = {
        "metric": "mean-field adjoint: loss + gradient time (774 nodes, 180 days, 30-day checkpoints)",
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": T,
        "warmup": 1,
        "ms_per_step": round(ms / T, 4),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"adjoint: national geography at scale {scale}, {len(rp.init_pop)} nodes, "
                               f"{int(rp.init_pop.sum())} agents aggregated at day 0, {T} days mean-field, ABM target"},
        "autograd_ms": round(t_ag * 1e3, 2),
        "rel_err_vs_autograd": {k: float(f"{v:.3e}") for k, v in rel_ag.items()},
        "rel_err_vs_central_fd": {k: float(f"{v:.3e}") for k, v in fd_err.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
