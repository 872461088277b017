"""Benchmark: agent-days/s of the laser-polio daily step on one MI355X.

Workload (the national-scale configuration): 774 nodes, populations
LogNormal(ln 2.5e5, 0.6) -> 2.27e8 agents, gravity network, RI + SIA campaigns,
2555 simulated days.  A "step" is one simulated day of the whole population.
The host initialises the population (not timed) and uploads it once.  The days
before the timed window run first (replayed hipGraphs), the last ``--warmup`` of
them as warm-up; then exactly ``--steps`` days starting at ``--start-day`` are
timed, bracketed by torch.cuda.synchronize(), with all state resident in HBM.
value = sum over the timed days of live agents / seconds.  The write-back of the
device-held agent records into the bound columns (lp_sync) is not part of a day
and is reported on its own (writeback_ms).

Also reported on the same JSON line:
  full_run      every simulated day of the run timed, as agent-days/s
  roofline      the dominant kernel, k_ei (DiseaseState over the E / I lists), on the
                window's days without VD / RI / SIA: algorithmic bytes (ei_pass_bytes
                below) over its HIP-event-timed duration against MI355X's 8 TB/s HBM
                peak; its HBM traffic per launch from the committed rocprofv3 record;
                whole_day: the counter bytes of every kernel of a window day over
                the day time; dense_byte_model: the bandwidth a dense day would need
  cpu_baseline  the CPU oracle (restated reference kernels, OpenMP on all host
                threads) on the same initial population for the first --cpu-days
                days, plus the next day on a single thread
  parity        the device state after that day (every results row, agent column
                and live end) against the oracle's, bit for bit; a mismatch makes
                the exit status 1
  cpu_baseline_spatial14   the CPU oracle over the whole 14-node workload
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


def ei_pass_bytes(res, t):
    """Algorithmic bytes of the DiseaseState pass (k_ei) of tick t: 17 B per agent exposed
    or infectious at the start of the day -- its exposure and infection timers read and
    written (2 x (2 + 2)), strain (1), daily_infectivity (4) and the paralysis fields (4).
    Susceptible and recovered agents cost nothing once the E and I agents are found
    without a scan (the E / I lists)."""
    return 17 * (int(res["E"][t - 1].astype(np.int64).sum()) + int(res["I"][t - 1].astype(np.int64).sum()))


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "day_kernels_national_traffic.md")


def measured_traffic(scale, tpw, path=TRAFFIC_PROFILE):
    """The committed rocprofv3 record of HBM bytes per dispatch of every day kernel
    (calibrated FETCH_SIZE and WRITE_SIZE; "kernels": averaged over the whole run,
    "window_kernels": over the dispatches of the days named by "window"), when it was
    measured on this population scale and span size; else None."""
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_per_kernel: (\{.*\})$", open(path).read(), re.M)
    if not m:
        return None
    rec = json.loads(m.group(1))
    if rec.get("workload") != f"scale={scale:g},tiles_per_block={tpw}":
        return None
    return rec


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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=float, default=1.0, help="population scale (1.0 = ~2.3e8 agents)")
    ap.add_argument("--start-day", type=int, default=1001,
                    help="first timed day (the days before it run untimed for the window, timed for the whole-run figure)")
    ap.add_argument("--cpu-days", type=int, default=24,
                    help="days of the CPU oracle baseline sample; the GPU state after the next day is compared with "
                         "the oracle's bit for bit (0 = skip both)")
    ap.add_argument("--cpu-spatial14", type=int, default=1, help="1: also time the CPU oracle on the 14-node workload")
    ap.add_argument("--dump-days", default="", help="write per-day N_live / N_EI / N_draw / flags to this CSV")
    ap.add_argument("--graph", type=int, default=1,
                    help="1: the timed window replays one hipGraph per day kind (as SEIR_ABM.run does); 0: it "
                         "launches every kernel from the host")
    ap.add_argument("--roofline-steps", type=int, default=20,
                    help="days after the timed window run with HIP events around every agent-pass launch")
    ap.add_argument("--roofline-pad-mib", type=int, default=4096,
                    help="MiB copied 4 times on the library's stream ahead of the roofline window")
    ap.add_argument("--tiles-per-block", type=int, default=0)
    for g in ("ei", "sample"):
        ap.add_argument(f"--{g}-grid", type=int, default=0, help=f"national: {g}_grid workgroups (0 = library default)")
    ap.add_argument("--init", choices=["device", "host"], default="device",
                    help="draw the initial population on the GPU (lp_init_population) or on the host (numpy)")
    ap.add_argument("--seeds", type=int, default=4, help="continental: replicate seeds batched in one context")
    ap.add_argument("--batch-scale", type=float, default=0.25, help="continental: population scale of the batched seeds")
    ap.add_argument("--workload", choices=["national", "replicates", "adjoint", "continental"], default="national",
                    help="national: one 774-node simulation (default); replicates: --sims independent "
                         "simulations of uneven sizes batched in one context, against serial runs of a sample; "
                         "adjoint: gradients of a mean-field calibration loss on the national geography")
    ap.add_argument("--sims", type=int, default=512, help="replicates: simulations in the batch")
    ap.add_argument("--mix-mode", type=int, default=0, choices=[0, 1],
                    help="replicates: 1 = the network mix of every simulation on MFMA (k_mix_mfma)")
    ap.add_argument("--serial-sample", type=int, default=16, help="replicates: simulations also run one at a time")
    ap.add_argument("--workers", type=int, default=16, help="replicates: host processes initialising simulations")
    ap.add_argument("--check-sims", type=int, default=8,
                    help="replicates: simulations of the batch replayed alone on the CPU oracle afterwards")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="instead of the benchmark: copy this many GiB device-to-device 5 times (a plain copy kernel "
                         "of known size, to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--trace-summary", default="",
                    help="instead of the benchmark: summarise a rocprofv3 kernel_trace.csv of it (run and window)")
    ap.add_argument("--traffic-summary", nargs=4, default=None, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="instead of the benchmark: the calibrated per-kernel HBM bytes of rocprofv3 "
                         "counter_collection.csv files (FETCH_SIZE, WRITE_SIZE of the benchmark, then of "
                         "--copy-calibration)")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30, help="bytes per copy of the calibration run")
    args = ap.parse_args()
    if args.trace_summary or args.traffic_summary:
        rocprof_summary(args)
        return
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
        print(f"bench.py: the library runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding the nodes over GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    if args.workload == "replicates":
        replicates(args)
        return
    if args.workload == "continental":
        continental(args)
        return
    if args.workload == "adjoint":
        adjoint_workload(args)
        return

    national(args)


RESULT_NAMES = [
    "S", "E", "I", "R", "E_by_strain", "I_by_strain", "potentially_paralyzed", "paralyzed",
    "new_potentially_paralyzed", "new_paralyzed", "pop", "births", "deaths", "new_exposed",
    "new_exposed_by_strain", "ri_vaccinated", "ri_protected", "ri_new_exposed_by_strain", "ipv_vaccinated",
    "sia_vaccinated", "sia_protected", "sia_new_exposed_by_strain",
]


def compare_with_oracle(sim, ref, t_last):
    """Differences between the device state after tick t_last (pulled to the host) and the
    oracle's: results rows 0..t_last of every array, every agent column over the whole
    capacity, every node's live end.  Empty when identical bit for bit."""
    bad = []
    for name in RESULT_NAMES:
        if sim.results.has(name):
            got = np.asarray(getattr(sim.results, name))[: t_last + 1]
            n = int(np.count_nonzero(got != ref.res[name][: t_last + 1]))
            if n:
                bad.append(f"results.{name}: {n} cells differ")
    for name, want in ref.cols.items():
        if sim.people.has(name):
            n = int(np.count_nonzero(getattr(sim.people, name) != want))
            if n:
                bad.append(f"people.{name}: {n} slots differ")
    if not np.array_equal(np.asarray(sim.people.live_end), ref.live_end):
        bad.append("live_end differs")
    return bad


def dense_model_bytes(res, flags, t0, t1):
    """Bytes per day of the dense day of the per-agent byte model (state + risk of every
    live agent, of every agent in a node with draws, 17 B per E/I agent, death dates on VD
    days, RI and SIA columns of every live agent), over ticks [t0, t1): the bandwidth a
    dense implementation would need to run the same days as fast."""
    from laser_polio_amd import engine as em

    pop_n = res["pop"].astype(np.int64)
    n_ei = res["E"].astype(np.int64).sum(axis=1) + res["I"].astype(np.int64).sum(axis=1)
    tx_new = res["new_exposed"].astype(np.int64) - res["ri_protected"] - res["sia_protected"]
    total = 0
    for t in range(t0, t1):
        n_live, f = int(pop_n[t].sum()), int(flags[t])
        b = 5 * n_live + 5 * int(pop_n[t][tx_new[t] > 0].sum()) + 17 * int(n_ei[t - 1])
        if f & em.DAY_VD_STEP:
            b += 4 * n_live
        if f & em.DAY_RI:
            b += 5 * n_live
        if f & em.DAY_SIA:
            b += 5 * n_live
        total += b
    return total


def _rocprof_days(path, value=False):
    """Rows of a rocprofv3 CSV (kernel trace, or counter collection with one counter) in
    dispatch order, each with the simulated tick it belongs to: the dispatches up to and
    including the t-th k_day_end make up tick t.  Returns [(tick, kernel name, ns or
    counter value)]; counter rows of one dispatch are summed."""
    import csv

    per = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            k = int(r["Dispatch_Id"])
            if value:
                v = float(r["Counter_Value"])
            else:
                v = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            if k in per:
                per[k][1] += v
            else:
                per[k] = [r["Kernel_Name"], v]
    out, tick = [], 0
    for k in sorted(per):
        name, v = per[k]
        out.append((tick, name, v))
        if "k_day_end" in name:
            tick += 1
    return out


def rocprof_summary(args):
    """--trace-summary: per-kernel mean duration over the whole run, the timed window and the
    roofline window of a rocprofv3 kernel trace of this benchmark; --traffic-summary:
    calibrated HBM bytes per dispatch from FETCH_SIZE / WRITE_SIZE runs of it, as the record
    that measured_traffic() reads."""
    import collections

    t0 = max(args.start_day, 1 + args.warmup)
    t1 = t0 + args.steps
    wins = {"run": (1, 1 << 30), "window": (t0, t1), "roofline_window": (t1, t1 + args.roofline_steps)}
    if args.trace_summary:
        rows = _rocprof_days(args.trace_summary)
        per = {w: collections.defaultdict(list) for w in wins}
        for tick, name, ns in rows:
            for w, (a, b) in wins.items():
                if a <= tick < b:
                    per[w][name].append(ns)
        head = " | ".join(f"dispatches, mean us ({w}" + (f" {a}-{b - 1})" if w != "run" else ")")
                          for w, (a, b) in wins.items())
        print(f"| kernel | {head} |")
        print("|---|" + "---:|" * len(wins))
        run = per["run"]
        for name in sorted(run, key=lambda n: -sum(run[n])):
            cells = []
            for w in wins:
                v = per[w].get(name, [])
                cells.append(f"{len(v)}, {sum(v) / len(v) / 1e3:.2f}" if v else "-")
            print(f"| `{name}` | " + " | ".join(cells) + " |")
        for w in ("window", "roofline_window"):
            days = (wins[w][1] - wins[w][0])
            print(f"\nsummed kernel time per day, {w}: {sum(map(sum, per[w].values())) / days / 1e3:.1f} us")
        return
    fetch, write, cfetch, cwrite = args.traffic_summary
    calib = []
    for path in (cfetch, cwrite):
        vals = [v for _, name, v in _rocprof_days(path, value=True) if "copy" in name.lower()]
        calib.append(max(vals) * 1024.0)         # KiB -> bytes of the largest copy dispatch
    n = args.copy_bytes
    f_scale, w_scale = n / calib[0], n / calib[1]
    rec = {"workload": f"scale={args.scale:g},tiles_per_block={args.tiles_per_block or 16}",
           "window": f"{t1}-{t1 + args.roofline_steps - 1}", "fetch_scale": round(f_scale, 4),
           "write_scale": round(w_scale, 4), "kernels": {}, "window_kernels": {}}
    for key, path, scale in (("fetch", fetch, f_scale), ("write", write, w_scale)):
        run, win = collections.defaultdict(list), collections.defaultdict(list)
        for tick, name, v in _rocprof_days(path, value=True):
            if tick >= 1:
                run[name].append(v * 1024.0 * scale)
            if t1 <= tick < t1 + args.roofline_steps:
                win[name].append(v * 1024.0 * scale)
        for dst, src in ((rec["kernels"], run), (rec["window_kernels"], win)):
            for name, v in src.items():
                e = dst.setdefault(name, {"dispatches": len(v)})
                e[key] = round(sum(v) / len(v))
    print("traffic_per_kernel: " + json.dumps(rec))


def national(args):
    import torch

    __graft_entry__.load_package()
    from laser_polio_amd import configs

    K, W = args.steps, args.warmup
    t_begin = max(args.start_day, 1 + W)
    t_end = t_begin + K
    t_init = time.perf_counter()
    pars = configs.nigeria_national_pars(scale=args.scale, dur=max(2555, t_end - 1))
    pars.init_on_device = args.init == "device"
    if args.tiles_per_block:
        pars.tiles_per_block = args.tiles_per_block
    for g in ("ei_grid", "sample_grid"):
        if getattr(args, g):
            pars[g] = getattr(args, g)
    sim = configs.build(pars)
    if pars.init_on_device:
        sim._ensure_engine()          # the population is drawn on the GPU here
        torch.cuda.synchronize()
    init_s = time.perf_counter() - t_init
    nt = sim.nt
    n0 = int(pars.init_pop.sum())
    log(f"[bench] init ({args.init}): {len(pars.init_pop)} nodes, {n0} agents, capacity {sim.people.capacity}, "
        f"{init_s:.2f}s")
    parity_day = args.cpu_days + 1 if args.cpu_days > 0 else 0
    if parity_day and parity_day >= t_begin - W:
        raise SystemExit(f"--cpu-days {args.cpu_days} must end before the warm-up days (start day {t_begin})")

    oracle_sim = None
    if parity_day:
        from oracle.oracle import OracleSim

        oracle_sim = OracleSim(sim.export_tables(), pop0=sim.pars.init_pop)

    t_up = time.perf_counter()
    eng = sim._ensure_engine()
    # the library works on a torch stream, so the roofline window can queue a device copy ahead of it
    stream = torch.cuda.Stream()
    eng.set_stream(stream)
    torch.cuda.synchronize()
    log(f"[bench] upload {time.perf_counter() - t_up:.1f}s")
    day_flags = eng.arrays["day_flags"]

    def timed_days(a, b, graph):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.run_days(a, b, use_graph=graph)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # tick 0 (initial tally and recount) is initialisation, not a simulated day
    eng.run_days(0, 1, use_graph=True)
    torch.cuda.synchronize()
    run_s = 0.0
    a = 1
    parity = None
    cpu = None
    if parity_day:
        run_s += timed_days(1, parity_day + 1, True)
        eng.sync()
        sim._mark_device_newer()
        from oracle import oracle as orc

        threads = int(orc.lib().orc_max_threads())
        oracle_sim.day(0)
        tc = time.perf_counter()
        for t in range(1, args.cpu_days + 1):
            oracle_sim.day(t)
        cpu_s = time.perf_counter() - tc
        cpu_ad = int(oracle_sim.res["pop"][1: args.cpu_days + 1].astype(np.int64).sum())
        oracle_sim.m.n_threads = 1
        tc = time.perf_counter()
        oracle_sim.day(parity_day)
        one_s = time.perf_counter() - tc
        one_ad = int(oracle_sim.res["pop"][parity_day].astype(np.int64).sum())
        cpu = {
            "value": round(cpu_ad / cpu_s, 1),
            "unit": "agent-days/s",
            "cores": threads,
            "kind": "port",
            "sample": f"days 1-{args.cpu_days} of the same {n0}-agent national population ({cpu_s:.1f}s, "
                      f"OpenMP {threads} threads, host nproc {os.cpu_count()})",
            "single_thread_value": round(one_ad / one_s, 1),
            "single_thread_sample": f"day {parity_day} on 1 thread ({one_s:.1f}s)",
        }
        log(f"[bench] cpu oracle: {cpu['value']:.4e} agent-days/s over {args.cpu_days} days ({cpu_s:.1f}s); "
            f"1 thread {cpu['single_thread_value']:.4e}")
        tc = time.perf_counter()
        bad = compare_with_oracle(sim, oracle_sim, parity_day)
        parity = {"oracle_ticks": f"0-{parity_day}", "identical": not bad, "mismatches": bad[:8],
                  "checked": "every results row, every agent column (whole capacity) and every node's live end"}
        log(f"[bench] parity vs oracle through day {parity_day}: {'identical' if not bad else bad[:4]} "
            f"({time.perf_counter() - tc:.1f}s)")
        del oracle_sim
        a = parity_day + 1
    if a < t_begin - W:
        run_s += timed_days(a, t_begin - W, True)
    run_s += timed_days(t_begin - W, t_begin, bool(args.graph))
    elapsed = timed_days(t_begin, t_end, bool(args.graph))
    run_s += elapsed
    # the roofline window: the next --roofline-steps days with every agent-pass launch
    # bracketed by HIP events on the library's stream (no graphs).  A device copy queued
    # ahead keeps the GPU busy while the host queues the window, so no bracket holds a
    # wait for the host.
    r0, r1 = t_end, min(nt, t_end + args.roofline_steps)
    launch_ms, launches, kernel_ms, rl_s = None, 0, 0.0, 0.0
    if r1 > r0:
        pad = torch.empty(args.roofline_pad_mib << 20, dtype=torch.uint8, device="cuda")
        pad_dst = torch.empty_like(pad)
        torch.cuda.synchronize()
        eng.set_timing(True)
        with torch.cuda.stream(stream):
            for _ in range(4):
                pad_dst.copy_(pad)
        eng.run_days(r0, r1, use_graph=False)
        torch.cuda.synchronize()
        launch_ms = eng.agent_pass_launch_ms()
        launches, kernel_ms = eng.agent_pass_timing()
        eng.set_timing(False)
        del pad, pad_dst
    if r1 < nt:
        run_s += timed_days(r1, nt, True)
    tc = time.perf_counter()
    eng.sync()     # write-back of the device-held records into the bound columns, and the error check
    sync_ms = (time.perf_counter() - tc) * 1e3
    sim._mark_device_newer()
    res = {k: np.asarray(getattr(sim.results, k)) for k in ("pop", "E", "I", "I_by_strain", "new_exposed",
                                                             "ri_protected", "sia_protected", "births", "deaths")}
    agent_days = int(res["pop"][t_begin:t_end].astype(np.int64).sum())
    value = agent_days / elapsed
    # the whole-run figure leaves out the roofline window (its launches are event-bracketed, not replayed)
    full_ad = int(res["pop"][1:nt].astype(np.int64).sum()) - int(res["pop"][r0:r1].astype(np.int64).sum())
    full_days = nt - 1 - (r1 - r0)
    log(f"[bench] {K} days ({t_begin}-{t_end - 1}) in {elapsed * 1e3:.3f} ms -> {value:.4e} agent-days/s; "
        f"whole run {full_days} days in {run_s:.3f}s -> {full_ad / run_s:.4e}; write-back {sync_ms:.2f} ms")
    if args.dump_days:
        dump_days(args.dump_days, res, day_flags, 1, nt)

    roofline = None
    ms_per_step = elapsed * 1e3 / K
    if launch_ms is not None and len(launch_ms) == r1 - r0:
        from laser_polio_amd import engine as em

        # the dominant kernel: k_ei, alone inside the bracket on days without VD / RI / SIA
        stage = em.DAY_VD_STEP | em.DAY_RI | em.DAY_SIA
        plain = [t for t in range(r0, r1) if not int(day_flags[t]) & stage]
        ei_bytes = sum(ei_pass_bytes(res, t) for t in plain)
        ei_ms = float(sum(launch_ms[t - r0] for t in plain))
        achieved = ei_bytes / (ei_ms / 1e3) / 1e9
        tpw = int(sim.pars.tiles_per_block or 16)
        rec = measured_traffic(args.scale, tpw)
        win = rec["window_kernels"] if rec and rec.get("window") == f"{r0}-{r1 - 1}" else None
        kei = "lp::k_ei(lp::Dev, int)"
        traffic = (win[kei]["fetch"] + win[kei]["write"]) if win and kei in win else None
        day_traffic = (sum((v["fetch"] + v["write"]) * v["dispatches"] for v in win.values()) / (r1 - r0)) if win else None
        net_bytes = 4 * len(pars.init_pop) ** 2
        roofline = {
            "bound": "hbm",
            "kernel": f"lp::k_ei (DiseaseState over the E / I lists; days {r0}-{r1 - 1} without VD / RI / SIA)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (round(traffic) if traffic else None),
            "launches": len(plain),
            "avg_launch_us": round(ei_ms * 1e3 / len(plain), 2),
            "bytes_per_launch": int(ei_bytes // len(plain)),
            "traffic_source": f"{os.path.relpath(TRAFFIC_PROFILE, ROOT)} (window {r0}-{r1 - 1})" if win else None,
            "agent_pass": {
                "launches": launches, "avg_us": round(kernel_ms * 1e3 / launches, 2),
                "note": "k_ei plus k_vd / k_ri_list / k_sia on VD / RI / SIA days, event-timed per day",
            },
            "whole_day": {
                "hbm_bytes_per_day": (int(day_traffic) if day_traffic else None),
                "ms_per_day": round(ms_per_step, 4),
                "achieved": (round(day_traffic / (ms_per_step / 1e3) / 1e9, 1) if day_traffic else None),
                "frac": (round(day_traffic / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if day_traffic else None),
                "days": f"counter bytes of days {r0}-{r1 - 1} over the day time of the timed window {t_begin}-{t_end - 1}",
                "network_bytes": net_bytes,
                "note": "counter bytes of every kernel of the window's days (rocprofv3 record) over the day time",
            },
            "dense_byte_model": {
                "bytes_per_day": int(dense_model_bytes(res, day_flags, t_begin, t_end) / K),
                "effective_GBs": round(dense_model_bytes(res, day_flags, t_begin, t_end) / elapsed / 1e9, 1),
                "note": "bandwidth a dense pass over every live agent would need to match this day time",
            },
        }
        if traffic:
            roofline["traffic_over_algorithmic"] = round(traffic / (ei_bytes / len(plain)), 3)
        log(f"[bench] k_ei: {len(plain)} plain days, avg {ei_ms * 1e3 / len(plain):.1f} us, {achieved:.0f} GB/s; "
            f"agent pass avg {kernel_ms * 1e3 / launches:.1f} us")

    c0 = None
    if args.cpu_spatial14:
        c0 = spatial14_cpu()

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "agent-days/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i8/u16/i32 state, f32 weights, int64 fixed-point sums",
        "data": "synthetic",
        "config": {
            "workload": f"national: 774 nodes, {n0} agents (population scale {args.scale}), RI + SIA every 180 d, "
                        f"{nt - 1} days; days {t_begin}-{t_end - 1} timed (days 1-{t_begin - 1} run before them, "
                        f"the last {W} as warm-up)",
            "nodes": int(len(pars.init_pop)),
            "agents": n0,
            "capacity": int(sim.people.capacity),
            "graph": bool(args.graph),
            "population_init": f"{args.init} ({init_s:.2f}s, not timed)",
        },
        "full_run": {"days": full_days, "seconds": round(run_s, 4), "agent_days_per_s": round(full_ad / run_s, 1),
                     "ms_per_day": round(run_s * 1e3 / full_days, 4),
                     "note": f"every simulated day of the run timed as replayed hipGraphs except the event-timed "
                             f"roofline window {r0}-{r1 - 1}; tick 0 and the oracle comparison excluded"},
        "writeback_ms": round(sync_ms, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_spatial14": c0,
        "parity": parity,
    }
    print(json.dumps(out), flush=True)
    if parity is not None and not parity["identical"]:
        sys.exit(1)


def spatial14_cpu():
    """The CPU oracle on the 14-node spatial workload (4.5e6 agents), all 365 days, on all
    host threads: the restated reference's agent-days/s at the size it runs on a CPU."""
    from laser_polio_amd import configs
    from oracle import oracle as orc
    from oracle.oracle import OracleSim

    pars = configs.spatial14_pars(scale=1.0)
    sim = configs.build(pars)
    o = OracleSim(sim.export_tables(), pop0=sim.pars.init_pop)
    o.day(0)
    tc = time.perf_counter()
    for t in range(1, sim.nt):
        o.day(t)
    s = time.perf_counter() - tc
    ad = int(o.res["pop"][1:].astype(np.int64).sum())
    threads = int(orc.lib().orc_max_threads())
    log(f"[bench] cpu oracle, 14-node workload: {ad / s:.4e} agent-days/s ({s:.1f}s, {threads} threads)")
    return {"value": round(ad / s, 1), "unit": "agent-days/s", "cores": threads, "kind": "port",
            "sample": f"14-node workload, {int(pars.init_pop.sum())} agents, days 1-{sim.nt - 1} ({s:.1f}s)"}


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
    from laser_polio_amd.batch import build_replicates, build_replicates_device

    S, dur = args.sims, 365
    pars_list = [configs.replicate_pars(i, scale=args.scale, dur=dur) for i in range(S)]
    for p in pars_list:
        p.mix_mode = args.mix_mode
    t_init = time.perf_counter()
    if args.init == "device":
        import torch

        batch = build_replicates_device(pars_list, configs.build, log=log)
    else:
        batch = build_replicates(pars_list, configs.build, workers=max(1, args.workers), log=log)
    n0 = int(sum(int(p.init_pop.sum()) for p in pars_list))
    log(f"[bench] replicates: {S} simulations, {batch.cfg.n_nodes} nodes, {n0} agents, capacity "
        f"{batch.cfg.capacity}, host construction {time.perf_counter() - t_init:.1f}s ({args.init} population)")

    import torch

    try:
        if args.tiles_per_block:
            tpw, tune = args.tiles_per_block, None
            batch.tiles_per_block = tpw
            batch.cfg.tiles_per_block = tpw
        else:
            tpw, tune = batch.autotune(grids=True)
            log(f"[bench] autotune tiles_per_block: " + ", ".join(f"{k}: {v * 1e3:.1f} ms" for k, v in tune.items())
                + f" -> {tpw}; grids " + ", ".join(
                    f"{g} {getattr(batch.cfg, g)} ({', '.join(f'{c}: {v * 1e3:.1f}' for c, v in batch.tune_log[g].items())} ms)"
                    for g in batch.GRID_CANDIDATES))
        sample = sorted(set(np.linspace(0, S - 1, min(args.serial_sample, S)).round().astype(int).tolist()))
        singles = [batch.subset([k], tiles_per_block=tpw) for k in sample]
        for one in singles:                 # the serial runs use the same launch shapes
            for g in batch.GRID_CANDIDATES:
                setattr(one.cfg, g, getattr(batch.cfg, g))
        t_up = time.perf_counter()
        eng = batch.prepare()
        torch.cuda.synchronize()
        log(f"[bench] upload / device population {time.perf_counter() - t_up:.1f}s")
        # the initial state of a sample of simulations, for their oracle replay afterwards
        check = sorted(set(np.linspace(0, S - 1, min(args.check_sims, S)).round().astype(int).tolist())) if args.check_sims else []
        if args.mix_mode == 1:
            check = []          # the MFMA mix is not a replay of the oracle's order
        init_state = {k: batch.pull_sim(k, names=None)[0] for k in check}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.run_days(0, batch.cfg.nt, use_graph=True)
        torch.cuda.synchronize()
        el_b = time.perf_counter() - t0
        eng.sync()
        ad_b = batch.agent_days(1, batch.cfg.nt)
        per_sim_ad = []
        pop = np.empty_like(batch.arrays["pop"])
        eng.pull("pop", pop)
        for k in range(S):
            per_sim_ad.append(int(pop[1:, batch.node_off[k]:batch.node_off[k + 1]].astype(np.int64).sum()))
        checks = batch_properties(batch, eng)
        log(f"[bench] batch properties: {checks}")
        parity = replicate_parity(batch, check, init_state, log) if check else None
        batch.close()
        log(f"[bench] batched: {S} sims x {dur} days in {el_b:.3f}s -> {S / el_b * 3600:.4e} sims/hr, "
            f"{ad_b / el_b:.4e} agent-days/s")

        times, ads = [], []
        for k, one in zip(sample, singles):
            e1 = one.prepare()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e1.run_days(0, one.cfg.nt, use_graph=True)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            e1.sync()
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
                   "tiles_per_block": int(tpw), "mix_mode": int(args.mix_mode),
                   "grids": {g: int(getattr(batch.cfg, g)) for g in batch.GRID_CANDIDATES}},
        "serial_sims_per_hr": round(S / serial_total * 3600, 1),
        "speedup_vs_serial": round(serial_total / el_b, 2),
        "batched_agent_days_per_s": round(ad_b / el_b, 1),
        "serial_sample": {"sims": len(times), "seconds": round(float(np.sum(times)), 4),
                          "fit_us_per_day": round(float(coef[0]) * 1e6, 2),
                          "fit_ps_per_agent_day": round(float(coef[1]) * 1e12, 3)},
        "batch_properties": checks,
        "parity": parity,
    }
    print(json.dumps(out), flush=True)
    if not checks["ok"] or (parity is not None and not parity["identical"]):
        sys.exit(1)


def batch_properties(batch, eng):
    """Bookkeeping every simulation of a finished batch must satisfy, from its results:
    pop[t] = pop[t-1] + births[t] - deaths[t] on every node and day, and S + E + I + R =
    pop on every node and logged day (through the simulation's early stop, if any)."""
    res = {}
    for name in ("pop", "births", "deaths", "S", "E", "I", "R"):
        out = np.empty_like(batch.arrays[name])
        eng.pull(name, out)
        res[name] = out.astype(np.int64)
    stops = eng.stop_days()
    bad_pop, bad_seir = 0, 0
    for k in range(len(batch.packs)):
        na, nb = batch.node_off[k], batch.node_off[k + 1]
        last = batch.cfg.nt - 1 if stops[k] < 0 else int(stops[k])
        r = {n: v[: last + 1, na:nb] for n, v in res.items()}
        bad_pop += int(np.count_nonzero(r["pop"][1:] != r["pop"][:-1] + r["births"][1:] - r["deaths"][1:]))
        bad_seir += int(np.count_nonzero(r["S"] + r["E"] + r["I"] + r["R"] != r["pop"]))
    return {"ok": bad_pop == 0 and bad_seir == 0, "pop_bookkeeping_mismatches": bad_pop,
            "seir_vs_pop_mismatches": bad_seir, "sims": len(batch.packs)}


CONTINENTAL_RESULTS = ("pop", "births", "deaths", "S", "E", "I", "R", "E_by_strain", "I_by_strain", "new_exposed",
                       "new_exposed_by_strain", "sia_new_exposed_by_strain")


def continental_single(scale):
    """One continental simulation at population `scale`, drawn on the GPU, every day timed
    as replayed hipGraphs; its results checked by result_properties and the state counts
    of 8 sampled nodes' agent columns against the last day's rows."""
    import torch

    __graft_entry__.load_package()
    from laser_polio_amd import configs

    t_init = time.perf_counter()
    pars = configs.continental_pars(scale=scale)
    pars.init_on_device = True
    sim = configs.build(pars)
    n0 = int(pars.init_pop.sum())
    eng = sim._ensure_engine()
    torch.cuda.synchronize()
    init_s = time.perf_counter() - t_init
    nt, p = sim.nt, sim.people
    log(f"[bench] continental: {len(pars.init_pop)} nodes, {n0} agents, capacity {p.capacity}, "
        f"built and drawn in {init_s:.1f}s")
    eng.run_days(0, 1, use_graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.run_days(1, nt, use_graph=True)
    torch.cuda.synchronize()
    run_s = time.perf_counter() - t0
    eng.sync()
    sim._mark_device_newer()
    res = {k: np.asarray(getattr(sim.results, k)) for k in CONTINENTAL_RESULTS}
    ad = int(res["pop"][1:nt].astype(np.int64).sum())
    props = result_properties(res)
    rng = np.random.default_rng(0)
    bad_nodes = []
    for n in sorted(rng.choice(len(pars.init_pop), size=8, replace=False).tolist()):
        a, b = int(p.seg_begin[n]), int(p.live_end[n])
        st = eng.tensors["disease_state"][a:b].cpu().numpy().view(np.int8)
        got = [int(np.count_nonzero(st == v)) for v in (0, 1, 2, 3)]
        want = [int(res[k][nt - 1, n]) for k in ("S", "E", "I", "R")]
        if got != want:
            bad_nodes.append(n)
    props["column_count_mismatch_nodes"] = bad_nodes
    props["ok"] = props["ok"] and not bad_nodes
    tx = (res["new_exposed_by_strain"] - res["sia_new_exposed_by_strain"]).astype(np.int64).sum(axis=(0, 1))
    log(f"[bench] continental: {nt - 1} days in {run_s:.3f}s -> {ad / run_s:.4e} agent-days/s; "
        f"transmission exposures by strain {tx.tolist()}; properties {props}")
    out = {"agents": n0, "capacity": int(p.capacity), "days": nt - 1, "seconds": run_s, "agent_days": ad,
           "init_s": init_s, "transmission_exposures_by_strain": tx.tolist(), "properties": props}
    eng.close()
    del sim, eng
    torch.cuda.empty_cache()
    return out


def continental_batched(seeds, scale):
    """`seeds` replicate seeds of the continental geography at population `scale`,
    batched in one library context (every kernel launch of a day covers all of them;
    their identical networks stored once, share_networks); each seed's results checked by
    result_properties."""
    import torch

    __graft_entry__.load_package()
    from laser_polio_amd import configs
    from laser_polio_amd.batch import build_replicates_device

    plist = [configs.continental_pars(scale=scale, run_seed=42 + k) for k in range(seeds)]
    tb0 = time.perf_counter()
    batch = build_replicates_device(plist, configs.build, share_networks=True)
    n_blocks = len(set(batch.net_off.tolist()))
    try:
        beng = batch.prepare()
        torch.cuda.synchronize()
        log(f"[bench] {seeds} seeds at scale {scale}: {batch.cfg.n_nodes} nodes, capacity {batch.cfg.capacity}, "
            f"{n_blocks} network block(s), "
            f"built and drawn in {time.perf_counter() - tb0:.1f}s")
        t0 = time.perf_counter()
        beng.run_days(0, batch.cfg.nt, use_graph=True)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        beng.sync()
        bres = {}
        for name in CONTINENTAL_RESULTS:
            out = np.empty_like(batch.arrays[name])
            beng.pull(name, out)
            bres[name] = out
        props = [result_properties({n: v[:, batch.node_off[k]:batch.node_off[k + 1]] for n, v in bres.items()})
                 for k in range(seeds)]
        ad = batch.agent_days(1, batch.cfg.nt)
    finally:
        batch.close()
        torch.cuda.empty_cache()
    out = {"seeds": seeds, "scale": scale, "agents": int(sum(int(q.init_pop.sum()) for q in plist)),
           "capacity": int(batch.cfg.capacity), "network_blocks": n_blocks, "seconds": round(el, 4), "agent_days_per_s": round(ad / el, 1),
           "properties_ok": all(q["ok"] for q in props), "properties": props}
    log(f"[bench] batched seeds: {out['agents']} agents, 365 days in {el:.3f}s -> {ad / el:.4e} agent-days/s; "
        f"properties ok: {out['properties_ok']}")
    return out


def continental(args):
    """The continental stress case (configs.continental_pars): 5000 nodes, about 1.36e9
    agents at scale 1, three strains (relative infectivity 1 / 0.25 / 0.125), SIA every
    120 days on 30 % of nodes alternating mOPV2 / nOPV2, births and deaths, 365 days: one
    simulation at `--scale` (continental_single), then `--seeds` replicate seeds of the
    same geography at `--batch-scale` batched in one context (continental_batched)."""
    one = continental_single(args.scale)
    batched = continental_batched(args.seeds, args.batch_scale) if args.seeds > 1 else None
    out = {
        "metric": "agent-days/s, continental stress case",
        "value": round(one["agent_days"] / one["seconds"], 1),
        "unit": "agent-days/s",
        "n_gpus": 1,
        "steps": one["days"],
        "warmup": 0,
        "ms_per_step": round(one["seconds"] * 1e3 / one["days"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i8/u16/i32 state, f32 weights, int64 fixed-point sums",
        "data": "synthetic",
        "config": {"workload": f"continental: 5000 nodes, {one['agents']} agents (scale {args.scale}), 3 strains, "
                               f"SIA every 120 d (mOPV2 / nOPV2), {one['days']} days", "capacity": one["capacity"],
                   "population_init": f"device ({one['init_s']:.1f}s with host construction, not timed)"},
        "transmission_exposures_by_strain": one["transmission_exposures_by_strain"],
        "properties": one["properties"],
        "batched_seeds": batched,
    }
    print(json.dumps(out), flush=True)
    if not one["properties"]["ok"] or (batched is not None and not batched["properties_ok"]):
        sys.exit(1)


def result_properties(res, last=None):
    """Bookkeeping one simulation's results must satisfy through tick `last` (default: all
    rows): pop[t] = pop[t-1] + births[t] - deaths[t] per node and day; S + E + I + R = pop;
    the per-strain E, I and new-exposure rows sum to their totals.  `res`: results arrays
    [t, node(, strain)].  Returns the mismatch counts."""
    r = {k: np.asarray(v, dtype=np.int64) for k, v in res.items()}
    if last is not None:
        r = {k: v[: last + 1] for k, v in r.items()}
    out = {
        "pop_bookkeeping_mismatches": int(np.count_nonzero(r["pop"][1:] != r["pop"][:-1] + r["births"][1:] - r["deaths"][1:])),
        "seir_vs_pop_mismatches": int(np.count_nonzero(r["S"] + r["E"] + r["I"] + r["R"] != r["pop"])),
        "strain_sum_mismatches": int(np.count_nonzero(r["E_by_strain"].sum(-1) != r["E"])
                                     + np.count_nonzero(r["I_by_strain"].sum(-1) != r["I"])
                                     + np.count_nonzero(r["new_exposed_by_strain"].sum(-1) != r["new_exposed"])),
    }
    out["ok"] = not any(out.values())
    return out


def replicate_parity(batch, ks, init_state, log):
    """Simulations ks of the batch, each replayed alone on the CPU oracle from its
    initial state: every results row, agent column and live end bit for bit."""
    from oracle.oracle import OracleSim
    from laser_polio_amd.model import AGENT_COLUMNS

    bad = []
    tc = time.perf_counter()
    for k in ks:
        pk = batch.packs[k]
        tb = pk.tables
        tb.cols = init_state[k]
        o = OracleSim(tb, pop0=pk.results["pop"][0]).run()
        tb.cols = None
        cols, res, live = batch.pull_sim(k)
        for name, want in o.res.items():
            if name in res and not np.array_equal(res[name], want):
                bad.append(f"sim {k} results.{name}")
        for name in AGENT_COLUMNS:
            if not np.array_equal(cols[name], o.cols[name]):
                bad.append(f"sim {k} people.{name}")
        if not np.array_equal(live, o.live_end):
            bad.append(f"sim {k} live_end")
    log(f"[bench] {len(ks)} sampled simulations replayed on the oracle: {'identical' if not bad else bad[:6]} "
        f"({time.perf_counter() - tc:.1f}s)")
    return {"sims": list(ks), "identical": not bad, "mismatches": bad[:8]}


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
    scale = 0.1 if args.scale == 1.0 else args.scale
    out = adjoint_result(scale, init=args.init)
    print(json.dumps(out), flush=True)


def adjoint_result(scale, init="device"):
    """The adjoint workload at population `scale` as a bench record (see adjoint_workload)."""
    import torch

    __graft_entry__.load_package()
    from laser_polio_amd import adjoint as adj
    from laser_polio_amd import configs

    T = 180
    pars = configs.nigeria_national_pars(scale=scale, dur=T)
    pars.init_on_device = init == "device"
    sim = configs.build(pars)
    rp = sim.pars
    log(f"[bench] adjoint: {len(rp.init_pop)} nodes, {int(rp.init_pop.sum())} agents (scale {scale}), {T} days")
    dev = torch.device("cuda", 0)
    prob = adj.MeanFieldProblem(sim, T, torch.float32, dev)     # compartments of day 0
    t0 = time.perf_counter()
    sim.run()
    y_daily = (sim.results.new_exposed.astype(np.int64) - sim.results.ri_protected - sim.results.sia_protected)[1:T + 1]
    log(f"[bench] ABM target: {int(y_daily.sum())} transmission exposures in {time.perf_counter() - t0:.1f}s")
    prob.set_target(y_daily)
    perturb = {"r0s": 0.9, "k": 0.8, "amp": 1.2}
    p32 = prob.params(**perturb)

    def grad_hip():
        p = {k: v.clone().requires_grad_(True) for k, v in p32.items()}
        L = prob.loss(p, adj.rollout)
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
    L_ag = prob.loss(p, adj.rollout_reference)
    L_ag.backward()
    torch.cuda.synchronize()
    t_ag = time.perf_counter() - t0
    rel_ag = {k: float((g_hip[k].double() - p[k].grad.double()).norm() / p[k].grad.double().norm().clamp_min(1e-30))
              for k in p32}
    prob.to(torch.float64, "cpu")
    p64 = prob.params(**perturb)
    f64 = lambda q: float(prob.loss(q, adj.rollout_reference))  # noqa: E731
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
    return out


if __name__ == "__main__":
    main()
