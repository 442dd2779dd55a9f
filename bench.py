#!/usr/bin/env python3
"""Throughput of SAMPO's per-iteration agent step on one MI355X.

A bench "step" is one simulated day: 24 reference iterations (one hour each,
src/abms.c:15-16) of kill -> update -> new eggs -> compaction over every agent
alive at the start of each iteration.  The unit of work is the agent-step:
one agent alive at the start of one iteration, processed through the whole
iteration; value = agent-steps per second of wall time over the timed days,
inputs resident in HBM.

Default workload: plain single run, large -- 32,000,000 initial agents (age
0-30 days, 50/50 sex, egg/larva/pupa/adult 30/30/10/30 %), buffer capacity
128,000,000, daily temperature 25 + 5 sin(2 pi d / 365) + N(0, 1.5) degC,
stage mortality 0.05/0.03/0.02/0.10 per day, one fused HIP pass per
iteration.

Batched workloads (--workload ensemble | sweep): every scenario of the set
is an instance of one engine, advanced by one launch of each kernel per
iteration and one HIP graph per day; value = agent-steps/s of the whole set
over the timed days.  The same line reports the plain paths on a uniform
sample of the set, one scenario at a time -- the fused engine alone, and the
reference's driver loop over the icl boundary -- and checks that each
sampled scenario's printStats rows are identical in the batched run, alone,
and through the driver loop.  Launch parameters come from the autotuner in
sampo_amd/batch.py (per initial-size bucket), run before the timed region.

Besides the JSON line's throughput, the single-run line carries:
  roofline     algorithmic bytes (34 B per agent in + 34 B per agent out) per
               fused-pass launch / that launch's duration from HIP events
               recorded around every pass of the timed iterations run again
               from the same inputs as one HIP graph with the events inside
               it (the timed run itself is the product's path, one day graph
               each with no events: an event pair costs an iteration about
               7 us, a third of a small run's iteration; both runs must give
               identical rows);
               traffic = HBM bytes per launch, the counter-measured ratio of
               traffic to algorithmic bytes (profiles/traffic_<workload>.json)
               times this run's algorithmic bytes per launch;
  cpu_baseline the C restatement of the reference's CPU-device path (oracle/:
               per-work-item kernels over OpenMP on every host core this
               process may use, the eight serialScan prefix counts of
               src/abms.c:416-439 single-threaded as the reference runs them),
               on a bounded sample of the same workload, plus a 1-thread
               figure;
  compaction_scan  the compaction index as a separate step (per-state ranks,
               include/sampo_compact.h) on as many synthetic keys as the
               workload's capacity: 5 bytes per agent / duration (HIP events).
The default (large) line also carries the other halves of the metric, each
as a sub-object measured in the same process:
  small_run    the 100k-agent single run (latency-bound: its own throughput and
               roofline fraction);
  batched      the 4096-scenario sweep in one engine against the plain
               one-scenario-at-a-time paths on a sample (the engine with its
               default launch parameters; the engine with the single-run
               policy: day launches and 512-agent chunks below 1M agents; the
               reference's driver loop), with the ratios and bit-identity
               counts;
  serving      a bounded window of the Poisson(50/s) serving trace: p50/p99
               completion latency and throughput.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "agent-steps/s batched vs plain on 1 GPU (bit-exact); %HBM peak; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
RECORD_BYTES = 34  # canonical agent record: 5 x f32 + 2 x u32 + 2 x u16 + 2 x u8

from sampo_amd.workloads import DMR, SINGLE_RUN as WORKLOADS  # noqa: E402


def build_inputs(w):
    from sampo_amd import workloads as W
    return W.single_run_inputs(w)


def traffic_ratio(name):
    """Measured HBM traffic over algorithmic bytes of the fused pass from the
    committed rocprofv3 counter summary (profiles/traffic_<workload>.json),
    or None."""
    p = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get("traffic_over_algorithmic")
    except (OSError, ValueError):
        return None


def host_cores():
    """Cores this process may run on: its CPU affinity, capped by a cgroup
    CPU quota if one is set."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if quota != "max":
            n = min(n, max(1, int(quota) // int(period)))
    except (OSError, ValueError):
        pass
    return n


def cpu_leg(w, days, scale, threads):
    """The oracle's CPU-device path (serialScan, src/abms.c:416-439) on the
    same generator at 1/scale of the population: (agent-steps/s, steps, s)."""
    import sampo_amd as S
    from oracle import pyoracle as O
    n0 = w["n0"] // scale
    temp = S.gen_temperature(w["seed"], days)
    layout, agents = S.gen_population(w["seed"], n0, float(temp[0]))
    sim = O.Sim(temp=temp, env=S.make_env(cc=w["cc_per_agent"] * n0), capacity=max(8 * n0, 4096),
                seed=w["seed"], layout=layout, agents=agents, mode=1, dmr=DMR, threads=threads, scan_path=1)
    t0 = time.perf_counter()
    rows = sim.run(days * 24)
    dt = time.perf_counter() - t0
    steps = int(rows["count"].astype(np.uint64).sum())
    return steps / dt, steps, dt, n0


def cpu_baseline(w, days, scale, scale1):
    cores = host_cores()
    v, steps, dt, n0 = cpu_leg(w, days, scale, cores)
    v1, steps1, dt1, n01 = cpu_leg(w, days, scale1, 1)
    return {"value": v, "unit": "agent-steps/s", "cores": cores, "kind": "port",
            "path": "CPU-device path: per-work-item kernels on OpenMP threads, eight single-threaded serialScan "
                    "prefix counts per iteration (src/abms.c:416-439)",
            "sample": f"first {days} day(s) of the same workload at 1/{scale} population ({n0} initial agents), "
                      f"{steps} agent-steps in {dt:.1f} s on {cores} threads",
            "one_thread": {"value": v1, "cores": 1,
                           "sample": f"first {days} day(s) at 1/{scale1} population ({n01} initial agents), "
                                     f"{steps1} agent-steps in {dt1:.1f} s"}}


TUNING_DIR = os.path.join(ROOT, "sampo_amd", "tuning")


def tuning_for(workload, args, specs, log):
    """Launch parameters per size bucket: the autotuner's choice for this
    workload kept in sampo_amd/tuning/<workload>.json (written by
    --retune), or a fresh autotuner run."""
    from sampo_amd import batch as B
    path = os.path.join(TUNING_DIR, f"{workload}.json")
    if args.no_tune:
        return B.Tuning(), {"source": "defaults"}
    if not args.retune and os.path.exists(path):
        return B.load_tuning(path), {"source": os.path.relpath(path, ROOT)}
    t0 = time.perf_counter()
    tun = B.autotune(specs, log=log)
    info = {"source": "autotuned in this run", "seconds": time.perf_counter() - t0}
    if args.retune:
        os.makedirs(TUNING_DIR, exist_ok=True)
        B.save_tuning(tun, path)
    return tun, info


def batched_measure(args, workload, log):
    """The batched set against the plain one-scenario-at-a-time paths on a
    uniform sample: (sub-object, everything bit-identical)."""
    from sampo_amd import batch as B
    from sampo_amd import workloads as W
    if workload == "ensemble":
        specs = W.ensemble(count=args.instances or 1024)
        desc = ("batched ensemble: replicates of the single-run parameters with distinct seeds, initial "
                "populations uniform 1e5-1e6, 365-day horizon, CC = 0.025 x N0")
    else:
        specs = W.sweep(count=args.instances or 4096)
        desc = ("uneven parameter sweep: initial population log-uniform 1e3-4e6, mean temperature 18-32 degC, "
                "amplitude 2-8 degC, per-stage mortality multipliers 0.5-1.5, 365-day horizon, CC = 0.0125 x N0")
    n_total = sum(sp.n0 for sp in specs)
    log(f"{workload}: {len(specs)} scenarios, {n_total / 1e9:.3f}e9 initial agents")
    tun, tinfo = tuning_for(workload, args, specs, log)
    log(f"tuning ({tinfo['source']}): tile per bucket {tun.tiles} (else {tun.tile}), chunk per bucket {tun.chunk}")
    sample = sorted(set(np.linspace(0, len(specs) - 1, min(args.plain_sample, len(specs))).round().astype(int)))
    bres, brows = B.run_batched(specs, tun, warm_days=args.batched_warmup, timed_days=args.batched_days,
                                keep_rows=sample, log=log)
    log(f"batched: {bres['agent_steps_per_s']:.3e} agent-steps/s, day {bres['day_ms']:.1f} ms")
    pres, prows = B.run_plain_engine([specs[i] for i in sample], days=1, warm_days=1)
    log(f"plain engine on {len(sample)} scenarios: {pres['agent_steps_per_s']:.3e} agent-steps/s")
    sres, srows = B.run_plain_engine([specs[i] for i in sample], days=1, warm_days=1, policy="single-run")
    log(f"plain engine, single-run policy: {sres['agent_steps_per_s']:.3e} agent-steps/s")
    isample = sample[::max(1, len(sample) // max(1, args.icl_sample))][:args.icl_sample]
    ires, irows = B.run_plain_icl([specs[i] for i in isample], days=1)
    log(f"plain icl driver on {len(isample)} scenarios: {ires['agent_steps_per_s']:.3e} agent-steps/s")
    n_cmp = 24 * 2 + 1  # the solo runs' warm-up and timed day
    ident = sum(B.same_rows(brows[i], r, n_cmp) for i, r in zip(sample, prows))
    ident_s = sum(B.same_rows(brows[i], r, n_cmp) for i, r in zip(sample, srows))
    ident_icl = sum(B.same_rows(brows[i], r, 24) for i, r in zip(isample, irows))
    sub = {
        "workload": desc, "instances": len(specs), "initial_agents": n_total,
        "agent_steps_per_s": bres["agent_steps_per_s"], "agent_days_per_s": bres["agent_steps_per_s"] / 24.0,
        "timed_days": args.batched_days, "warmup_days": args.batched_warmup, "day_ms": bres["day_ms"],
        "plain_engine": pres, "plain_engine_single_run_policy": sres, "plain_icl": ires,
        "speedup_vs_plain_engine": bres["agent_steps_per_s"] / pres["agent_steps_per_s"],
        "speedup_vs_plain_engine_single_run_policy": bres["agent_steps_per_s"] / sres["agent_steps_per_s"],
        "speedup_vs_plain_icl": bres["agent_steps_per_s"] / ires["agent_steps_per_s"],
        "identity": {"checked": len(sample), "identical_to_solo": int(ident),
                     "identical_to_solo_single_run_policy": int(ident_s), "rows_compared": n_cmp,
                     "checked_icl": len(isample), "identical_to_icl_driver": int(ident_icl),
                     "rows_compared_icl": 24},
        "roofline": {"bound": "hbm", "achieved": bres["pass_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bres["pass_GBs"] / HBM_PEAK_GBS, "kernel": "k_pass",
                     "bytes": "34 B per agent in + 34 B per agent out of each iteration, HIP events over one day"},
        "tile": "x".join(map(str, tun.tile)), "chunk_per_bucket": {str(k): v for k, v in sorted(tun.chunk.items())},
        "tile_per_bucket": {str(k): "x".join(map(str, v)) for k, v in sorted(tun.tiles.items())},
        "tuning": {**tinfo, **tun.to_json()},
        "detail": bres,
    }
    return sub, ident == len(sample) and ident_s == len(sample) and ident_icl == len(isample)


def batched_main(args):
    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    sub, ok = batched_measure(args, args.workload, log)
    out = {
        "metric": METRIC, "value": sub["agent_steps_per_s"], "unit": "agent-steps/s", "n_gpus": 1,
        "steps": args.batched_days, "warmup": args.batched_warmup, "ms_per_step": sub["day_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+u32", "data": "synthetic",
        "config": {"workload": sub["workload"], "instances": sub["instances"], "initial_agents": sub["initial_agents"],
                   "timed_days": args.batched_days, "iterations_per_day": 24, "tile": sub["tile"],
                   "chunk_per_bucket": sub["chunk_per_bucket"], "parallelism": "1 GPU"},
        "roofline": {**sub["roofline"], "traffic": None},
        "batched": sub,
    }
    print(json.dumps(out), flush=True)
    if not ok:
        sys.exit("a sampled scenario differs between the batched run and its run alone or the driver loop")


def serve_measure(args, window, log):
    from sampo_amd import serve as V
    from sampo_amd import workloads as W
    specs = W.serving(rate=args.rate, duration=window)
    offered = sum(sp.n0 * sp.horizon for sp in specs)
    log(f"serve: {len(specs)} requests over {window:.0f} s at {args.rate:.0f}/s, "
        f"{offered / 1e12:.2f}e12 initial-size agent-steps offered")
    srv = V.Server(max_slots=args.slots, log=log, policy=args.policy, min_active=args.min_active)
    results, summ = srv.run(specs, clock="wall")
    sub = {"workload": "serving: Poisson arrivals, initial population log-uniform 1e4-1e7, horizon 30-365 days, "
                       "temperature/mortality as the sweep, CC = 0.0125 x N0",
           "rate_per_s": args.rate, "arrival_window_s": window, "requests": len(specs), "slots": args.slots,
           "policy": args.policy, "min_active_agents": args.min_active,
           "latency_s": {"p50": summ.get("p50"), "p99": summ.get("p99"), "mean": summ.get("mean"),
                         "max": summ.get("max")},
           "agent_steps_per_s": summ["agent_steps_per_s"], "offered_initial_size_agent_steps": offered,
           "summary": summ}
    done = summ["completed"] + summ["rejected"] + summ.get("growth_failed", 0)
    return sub, done == len(specs)


def serve_main(args):
    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    sub, ok = serve_measure(args, args.window, log)
    summ = sub["summary"]
    out = {
        "metric": METRIC, "value": sub["agent_steps_per_s"], "unit": "agent-steps/s", "n_gpus": 1,
        "steps": summ["days_stepped"], "warmup": 0, "ms_per_step": summ["wall_s"] * 1e3 / max(1, summ["days_stepped"]),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+u32", "data": "synthetic",
        "config": {"workload": sub["workload"], "rate_per_s": args.rate, "arrival_window_s": args.window,
                   "requests": sub["requests"], "slots": args.slots, "policy": args.policy,
                   "min_active_agents": args.min_active, "parallelism": "1 GPU"},
        "latency_s": sub["latency_s"],
        "serving": sub,
    }
    print(json.dumps(out), flush=True)
    if not ok:
        sys.exit("not every request completed")


def single_run(w, args, name, steps, warmup, chunk, day=0):
    """One scenario in its own engine: (sub-object with throughput and
    roofline, inputs resident).  The scenario runs twice from the same
    inputs: the timed days as the product runs them (one day graph each, or
    with day > 0 one cooperative day launch of `day` work-groups each, no
    events), then the same days again with HIP events: around each pass of
    the iterations replayed as one graph, or around each day launch.  Both
    runs must give identical rows."""
    import sampo_amd as S
    temp, env, layout, agents = build_inputs(w)
    wg, apt, waves = (int(x) for x in (args.tile.split("x") + ["0"])[:3])
    first = 24 * warmup
    n_iter = 24 * steps

    def run(timed):
        with S.Engine(max_slots=1) as e:
            e.set_tile(wg, apt, waves)
            e.set_grid(args.per_cu)
            e.set_chunk(chunk, args.target_chunks)
            e.set_day(day)
            t_admit = time.perf_counter()
            slot = e.admit(mode=S.MODE_CONFIG, seed=w["seed"], horizon=24 * w["days"], temp=temp, env=env,
                           max_agents=w["max_agents"], dmr=DMR, layout=layout, agents=agents)
            e.sync()
            admit_ms = (time.perf_counter() - t_admit) * 1e3
            if warmup:
                e.step(24 * warmup)
            out = {"admit_ms": admit_ms}
            if timed:
                e.step_days(0)  # the day graph, captured before the timed region
                e.sync()
                e.agent_steps(reset=True)
                t0 = time.perf_counter()
                e.step_days(steps)  # the product's path: one day graph per day
                e.sync()
                out["dt"] = time.perf_counter() - t0
                out["steps_done"] = e.agent_steps(reset=True)
            elif day:
                out["wall_ms"], out["pass_ms"] = e.time_days(steps)
            else:
                e.capture(n_iter)
                e.sync()
                e.replay()
                e.sync()
                out["wall_ms"], out["pass_ms"] = e.replay_times()
            status, step, _ = e.status(slot)
            if status not in (S.RUNNING, S.DONE) or step != first + n_iter:
                sys.exit(f"instance stopped early: status {status} at iteration {step}")
            out["rows"] = e.rows(slot, first, n_iter + 1)
        return out

    a = run(timed=True)
    b = run(timed=False)
    del agents
    if not np.array_equal(a["rows"], b["rows"]):
        sys.exit("the timed run and its replay with events differ")
    dt, steps_done, rows = a["dt"], a["steps_done"], a["rows"]
    wall_ms, pass_ms, admit_ms = b["wall_ms"], b["pass_ms"], a["admit_ms"]
    counts = rows["count"].astype(np.uint64).sum(axis=1)
    assert int(counts[:-1].sum()) == steps_done
    launches = steps if day else n_iter
    per_launch = RECORD_BYTES * (int(counts[:-1].sum()) + int(counts[1:].sum())) / launches
    achieved = per_launch / (pass_ms / launches / 1e3) / 1e9
    ratio = None if day else traffic_ratio(name)
    value = steps_done / dt
    return {
        "value": value, "steps": steps, "warmup": warmup, "ms_per_step": dt * 1e3 / steps,
        "config": {"workload": w["desc"], "initial_agents": w["n0"], "max_agents": w["max_agents"],
                   "carrying_capacity": w["cc_per_agent"] * w["n0"], "timed_days": steps,
                   "iterations_per_day": 24, "tile": "512x1 (day launch)" if day else args.tile,
                   "day_launch_work_groups": day, "chunk": chunk,
                   "target_chunks": args.target_chunks, "agents_at_end": int(counts[-1]),
                   "mean_agents": steps_done / n_iter, "parallelism": "1 GPU"},
        "agent_days_per_s": value / 24.0,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None if ratio is None else ratio * per_launch,
                     "traffic_over_algorithmic": ratio,
                     # the bytes the pass really moves (the record is stored in fewer bytes than the 34-B
                     # algorithmic one): their rate as a fraction of the peak
                     "traffic_frac": None if ratio is None else achieved * ratio / HBM_PEAK_GBS,
                     "kernel": "k_pass<512, 1, 0, true> (day launch: 24 iterations, passes and folds)" if day
                               else "k_pass", "launches": launches,
                     "measured_over": (f"the {steps} timed days, run again from the same inputs (identical rows), "
                                       "an event pair around each day launch") if day else
                                      (f"the {n_iter} timed iterations, run again from the same inputs (identical "
                                       "rows) as one HIP graph with an event pair around each pass"),
                     "algorithmic_bytes_per_launch": per_launch, "mean_launch_ms": pass_ms / launches},
        "roofline_replay_wall_ms": wall_ms,
        "admit_ms": admit_ms,
    }


def day_groups(args, w):
    """Work-groups of the day launches of a single run: --day, or 256 (one
    per CU) below 1M initial agents, where a launch per iteration costs more
    than the agents; 0 (per-iteration launches) above."""
    if args.day >= 0:
        return args.day
    return 256 if w["n0"] < 1_000_000 else 0


def per_iteration_summary(r):
    """The figures of a single run by per-iteration launches (day graphs)."""
    return {"value": r["value"], "ms_per_step": r["ms_per_step"], "agent_days_per_s": r["agent_days_per_s"],
            "roofline_frac": r["roofline"]["frac"], "mean_pass_ms": r["roofline"]["mean_launch_ms"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed days")
    ap.add_argument("--warmup", type=int, default=2, help="untimed days before the timed region")
    ap.add_argument("--workload", default="large", choices=sorted(WORKLOADS) + ["ensemble", "sweep", "serve"])
    ap.add_argument("--no-extras", action="store_true",
                    help="large: only the single run (no small_run, batched and serving sub-objects)")
    ap.add_argument("--instances", type=int, default=0, help="batched workloads: scenarios (0: the full set)")
    ap.add_argument("--batched-days", type=int, default=7, help="batched workloads: timed days")
    ap.add_argument("--batched-warmup", type=int, default=1, help="batched workloads: untimed days first")
    ap.add_argument("--plain-sample", type=int, default=64, help="batched workloads: scenarios run alone")
    ap.add_argument("--icl-sample", type=int, default=16, help="batched workloads: scenarios through the driver loop")
    ap.add_argument("--no-tune", action="store_true", help="batched workloads: default launch parameters")
    ap.add_argument("--retune", action="store_true",
                    help="batched workloads: run the autotuner and keep its choice in sampo_amd/tuning/")
    ap.add_argument("--rate", type=float, default=50.0, help="serve: arrivals per second")
    ap.add_argument("--window", type=float, default=10.0, help="serve: seconds of arrivals")
    ap.add_argument("--extra-window", type=float, default=4.0, help="large: seconds of arrivals of the serving part")
    ap.add_argument("--slots", type=int, default=2048, help="serve: engine slots")
    ap.add_argument("--policy", default="srpt", choices=["srpt", "all"],
                    help="serve: advance the requests with the least work left first, or all every day")
    ap.add_argument("--min-active", type=int, default=64_000_000,
                    help="serve, srpt: agents kept in each day's step at least")
    ap.add_argument("--tile", default="512x1x6", help="work-group size x agents per thread x min waves per SIMD")
    ap.add_argument("--per-cu", type=int, default=0, help="persistent work-groups of the pass per CU (0: all that fit)")
    ap.add_argument("--chunk", type=int, default=0,
                    help="single run: minimum agents per chunk (0: 3072, or 512 below 1M initial agents, where "
                         "3072-agent chunks would leave most work-groups idle)")
    ap.add_argument("--target-chunks", type=int, default=0,
                    help="single run: most chunks per iteration (0: two per resident work-group)")
    ap.add_argument("--day", type=int, default=-1,
                    help="single runs: work-groups of the cooperative day launches (0: per-iteration launches; "
                         "default: 256 below 1M initial agents, else 0)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-days", type=int, default=3, help="CPU baseline: days simulated")
    ap.add_argument("--cpu-scale", type=int, default=0,
                    help="CPU baseline: population divisor of the all-core leg (0: 10 for the large run, 1 small)")
    ap.add_argument("--cpu-scale1", type=int, default=0,
                    help="CPU baseline: population divisor of the 1-thread leg (0: 20 large, 2 small)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path runs on one GPU, "
              "multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    if args.workload in ("ensemble", "sweep"):
        return batched_main(args)
    if args.workload == "serve":
        return serve_main(args)
    import sampo_amd as S

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    w = WORKLOADS[args.workload]
    if args.warmup + args.steps > w["days"]:
        sys.exit("warmup + steps exceeds the workload's horizon")
    chunk = args.chunk or (512 if w["n0"] < 1_000_000 else 3072)
    day = day_groups(args, w)
    res = single_run(w, args, args.workload, args.steps, args.warmup, chunk, day)
    if day:  # the same run by per-iteration launches, for comparison
        res["per_iteration_launches"] = per_iteration_summary(single_run(w, args, args.workload, args.steps,
                                                                         args.warmup, chunk, 0))
    log(f"{args.workload}: {res['value']:.3e} agent-steps/s, frac {res['roofline']['frac']:.3f}")
    copy_bytes = 1 << 30
    copy_ms = S.stream_copy_ms(copy_bytes)
    res["roofline"]["stream_copy_GBs"] = 2 * copy_bytes / (copy_ms / 1e3) / 1e9
    # the compaction index as a separate step (include/sampo_compact.h): a
    # 1-byte key in and a 4-byte rank out per agent, at the workload's capacity
    scan_n = int(w["max_agents"])
    scan_ms = S.compact_time(scan_n, iters=20, seed=w["seed"])
    out = {"metric": METRIC, "value": res["value"], "unit": "agent-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+u32", "data": "synthetic"}
    out.update({k: v for k, v in res.items() if k not in ("value", "steps", "warmup", "ms_per_step")})
    out["compaction_scan"] = {"agents": scan_n, "bytes_per_agent": 5, "ms": scan_ms,
                              "GBs": 5 * scan_n / (scan_ms / 1e3) / 1e9,
                              "frac": 5 * scan_n / (scan_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                              "kernels": "k_count, k_scan_local, k_scan_top, k_rank"}
    ok = True
    if args.workload == "large" and not args.no_extras:
        sw = WORKLOADS["small"]
        small = single_run(sw, args, "small", 20, 2, 512, day_groups(args, sw))
        if small["config"]["day_launch_work_groups"]:
            small["per_iteration_launches"] = per_iteration_summary(single_run(sw, args, "small", 20, 2, 512, 0))
        log(f"small: {small['value']:.3e} agent-steps/s, frac {small['roofline']['frac']:.3f}")
        out["small_run"] = small
        out["batched"], ok_b = batched_measure(args, "sweep", log)
        out["serving"], ok_s = serve_measure(args, args.extra_window, log)
        ok = ok_b and ok_s
    if not args.no_cpu:
        small_w = args.workload == "small"
        scale = args.cpu_scale or (1 if small_w else 10)
        scale1 = args.cpu_scale1 or (2 if small_w else 20)
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_days, scale, scale1)
    print(json.dumps(out), flush=True)
    if not ok:
        sys.exit("a batched or served scenario differs from its run alone, or a request did not complete")


if __name__ == "__main__":
    main()
