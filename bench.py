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

Besides the JSON line's throughput, the line carries:
  roofline     algorithmic bytes (34 B per agent in + 34 B per agent out) per
               fused-pass launch / that launch's duration from HIP events
               recorded around every pass of the timed region;
  cpu_baseline the C restatement of the reference (oracle/), OpenMP over all
               host cores, on a bounded sample of the same workload.
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

WORKLOADS = {
    # configs[1] of the benchmark definition
    "large": dict(
        desc="plain single run, large: 32M initial agents, buffer capacity 128M, 365-day horizon, "
             "config-mode mortality 0.05/0.03/0.02/0.1 per day, T(d)=25+5sin(2pi d/365)+N(0,1.5)",
        n0=32_000_000, max_agents=128_000_000, cc_per_agent=0.025, seed=1, days=365),
    # configs[0]
    "small": dict(
        desc="plain single run, small: 100k initial agents, 365-day horizon, config-mode mortality",
        n0=100_000, max_agents=16_000_000, cc_per_agent=0.025, seed=1, days=365),
}
DMR = (0.05, 0.03, 0.02, 0.1)


def build_inputs(w):
    import sampo_amd as S
    temp = S.gen_temperature(w["seed"], w["days"])
    layout, agents = S.gen_population(w["seed"], w["n0"], float(temp[0]))
    env = S.make_env(cc=w["cc_per_agent"] * w["n0"])
    return temp, env, layout, agents


def traffic_from_profiles(name):
    """HBM bytes per fused-pass launch from the committed rocprofv3 counter
    summary (profiles/traffic_<workload>.json), or None."""
    p = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(w, days, scale):
    """The oracle (C restatement of the reference, gcc -O2, OpenMP over the
    host cores) on the same generator at 1/scale of the population."""
    import sampo_amd as S
    from oracle import pyoracle as O
    n0 = w["n0"] // scale
    temp = S.gen_temperature(w["seed"], days)
    layout, agents = S.gen_population(w["seed"], n0, float(temp[0]))
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    sim = O.Sim(temp=temp, env=S.make_env(cc=w["cc_per_agent"] * n0), capacity=max(8 * n0, 4096),
                seed=w["seed"], layout=layout, agents=agents, mode=1, dmr=DMR, threads=cores)
    t0 = time.perf_counter()
    rows = sim.run(days * 24)
    dt = time.perf_counter() - t0
    steps = int(rows["count"].astype(np.uint64).sum())
    return {"value": steps / dt, "unit": "agent-steps/s", "cores": cores, "kind": "port",
            "sample": f"first {days} day(s) of the same workload at 1/{scale} population ({n0} initial agents), "
                      f"{steps} agent-steps in {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed days")
    ap.add_argument("--warmup", type=int, default=2, help="untimed days before the timed region")
    ap.add_argument("--workload", default="large", choices=sorted(WORKLOADS))
    ap.add_argument("--tile", default="512x1x6", help="work-group size x agents per thread x min waves per SIMD")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-days", type=int, default=2)
    ap.add_argument("--cpu-scale", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path runs on one GPU, "
              "multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import sampo_amd as S
    w = WORKLOADS[args.workload]
    if args.warmup + args.steps > w["days"]:
        sys.exit("warmup + steps exceeds the workload's horizon")
    temp, env, layout, agents = build_inputs(w)
    wg, apt, waves = (int(x) for x in (args.tile.split("x") + ["0"])[:3])
    with S.Engine(max_slots=1) as e:
        e.set_tile(wg, apt, waves)
        t_admit = time.perf_counter()
        slot = e.admit(mode=S.MODE_CONFIG, seed=w["seed"], horizon=24 * w["days"], temp=temp, env=env,
                       max_agents=w["max_agents"], dmr=DMR, layout=layout, agents=agents)
        e.sync()
        admit_ms = (time.perf_counter() - t_admit) * 1e3
        del agents
        if args.warmup:
            e.step(24 * args.warmup)
        e.sync()
        e.agent_steps(reset=True)
        first = 24 * args.warmup
        n_iter = 24 * args.steps
        t0 = time.perf_counter()
        wall_ms, pass_ms = e.time_pass(n_iter)
        e.sync()
        dt = time.perf_counter() - t0
        status, step, n_now = e.status(slot)
        if status not in (S.RUNNING, S.DONE) or step != first + n_iter:
            sys.exit(f"instance stopped early: status {status} at iteration {step}")
        steps_done = e.agent_steps()
        rows = e.rows(slot, first, n_iter + 1)
        counts = rows["count"].astype(np.uint64).sum(axis=1)
        assert int(counts[:-1].sum()) == steps_done
        alg_bytes = RECORD_BYTES * (int(counts[:-1].sum()) + int(counts[1:].sum()))
        per_launch = alg_bytes / n_iter
        achieved = per_launch / (pass_ms / n_iter / 1e3) / 1e9
        copy_bytes = 1 << 30
        copy_ms = S.stream_copy_ms(copy_bytes)
    value = steps_done / dt
    out = {
        "metric": METRIC, "value": value, "unit": "agent-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+u32", "data": "synthetic",
        "config": {"workload": w["desc"], "initial_agents": w["n0"], "max_agents": w["max_agents"],
                   "carrying_capacity": w["cc_per_agent"] * w["n0"], "timed_days": args.steps,
                   "iterations_per_day": 24, "tile": args.tile, "agents_at_end": n_now,
                   "mean_agents": steps_done / n_iter, "parallelism": "1 GPU"},
        "agent_days_per_s": value / 24.0,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(args.workload),
                     "kernel": "k_pass", "launches": n_iter, "algorithmic_bytes_per_launch": per_launch,
                     "mean_launch_ms": pass_ms / n_iter, "stream_copy_GBs": 2 * copy_bytes / (copy_ms / 1e3) / 1e9},
        "gpu_wall_ms_events": wall_ms,
        "admit_ms": admit_ms,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_days, args.cpu_scale)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
