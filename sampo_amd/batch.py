"""Batched runs of many scenarios, the plain one-scenario-at-a-time paths
they are measured against, and the autotuner that picks launch parameters
per initial-size bucket.

Batched: every scenario is an instance of one engine; each iteration is one
k_plan + one persistent k_pass + one k_prep launch over all of them, and a
day is one HIP graph of 24 such iterations.  Between days the host polls the
instances and grows those that outgrow their buffers
(sampo_engine_service).

Plain paths, one scenario at a time:
  engine  the same fused kernels with one instance per engine (the
          single-run configuration of the large workload);
  icl     the reference's driver loop over the icl boundary
          (sampo_abms_run): the reference's kernel sequence and launch
          widths, a blocking read of the population every iteration, AoS
          buffers sized to the scenario's capacity.

Autotuner: initial sizes fall in buckets of a factor of 4 from 1e3.  A
uniform sample of the set large enough to fill the GPU (~3e8 agents) is
admitted for each trial and timed over one day after a warm-up day.  Two
rounds of coordinate descent, each visiting the buckets largest share of
agents first: for a bucket, its instances' value is set to each candidate
with the other buckets' kept, and the fastest in agent-steps/s is kept.
The first round tunes the chunk size (agents per work unit of the pass and
per output run), the second the tile shape of the pass (work-group size,
agents per thread, register cap).  Instances of different tile shapes are
advanced by one pass launch per shape, so a bucket's shape is its own; the
engine runs at most 4 shapes at once, and candidates that would need a
fifth are skipped.  Instances of a bucket are admitted with its chunk size
and tile shape.
"""
import json
import math
import time
from dataclasses import dataclass, field

import numpy as np

import sampo_amd as S
from sampo_amd import workloads as W

BUCKET_BASE = 1e3
CHUNK_CANDIDATES = (4096, 16384, 65536, 262144)
TILE_CANDIDATES = ((512, 1, 6), (768, 1, 6), (256, 1, 8), (1024, 1, 6), (512, 2, 0), (256, 2, 6))
MAX_TILE_SHAPES = 4  # pass launches per iteration the engine keeps (MAX_CLASSES)
DEFAULT_CHUNK = 16384
NO_CHUNK_CAP = 1 << 24  # target chunks per instance: effectively unlimited


def bucket_of(n0):
    return max(0, int(math.floor(math.log(max(n0, 1) / BUCKET_BASE, 4))))


def bucket_range(b):
    return BUCKET_BASE * 4 ** b, BUCKET_BASE * 4 ** (b + 1)


@dataclass
class Tuning:
    tile: tuple = (512, 1, 6)                    # shape of buckets without one of their own
    chunk: dict = field(default_factory=dict)   # bucket -> agents per chunk
    tiles: dict = field(default_factory=dict)   # bucket -> tile shape
    table: list = field(default_factory=list)   # measurements

    def chunk_for(self, n0):
        return self.chunk.get(bucket_of(n0), DEFAULT_CHUNK)

    def tile_for(self, n0):
        return tuple(self.tiles.get(bucket_of(n0), self.tile))

    def to_json(self):
        return {"tile": list(self.tile), "chunk": {str(k): v for k, v in sorted(self.chunk.items())},
                "tiles": {str(k): list(v) for k, v in sorted(self.tiles.items())}, "table": self.table}


def admit_specs(engine, specs, tuning, phys_factor=0.0, kws=None):
    """Admit specs in order with the tuned chunk size and tile shape of each
    one's bucket; returns the slots.  kws: already materialised admit kwargs
    (else generated on a thread pool)."""
    slots = []
    src = zip(specs, kws) if kws is not None else W.materialise_all(specs, phys_factor)
    for sp, kw in src:
        engine.set_tile(*tuning.tile_for(sp.n0))
        engine.set_chunk(tuning.chunk_for(sp.n0), NO_CHUNK_CAP)
        slots.append(engine.admit(**kw))
    return slots


def coordinate_descent(order, candidates, state, measure, allowed=lambda state: True, log=None, what=""):
    """One round of coordinate descent: for each key of `order` in turn, try
    every candidate value for it with the other keys kept, and keep the
    value with the highest measure(state).  Trials whose state fails
    allowed() are skipped.  Returns (state, best, table of trials)."""
    state = dict(state)
    best = measure(state)
    table = [{"start": True, "state": {str(k): v for k, v in state.items()}, "rate": best}]
    for key in order:
        for c in candidates:
            if c == state[key]:
                continue
            trial = {**state, key: c}
            if not allowed(trial):
                continue
            rate = measure(trial)
            table.append({"key": key, "value": c, "rate": rate})
            if log:
                log(f"autotune {what} bucket {key} = {c}: {rate:.3e} agent-steps/s")
            if rate > best:
                best, state = rate, trial
    return state, best, table


def _day_rate(engine, days=1):
    """Agent-steps per second over `days` graph-launched days (one warm-up
    day first builds the graph and fills the caches)."""
    engine.step_days(1)
    engine.service(1.25)
    engine.agent_steps(reset=True)
    ms = engine.time(24 * days, graph=True)
    return engine.agent_steps() / (ms / 1e3), ms


def tuning_mix(specs, agents=300_000_000, max_instances=1024):
    """A uniform sample of the set with about `agents` initial agents --
    enough to fill the GPU, so that each candidate is timed in the regime
    the full batch runs in."""
    total = sum(sp.n0 for sp in specs)
    k = max(1, int(math.ceil(total / agents)))
    mix = specs[::k]
    if len(mix) > max_instances:
        mix = mix[::int(math.ceil(len(mix) / max_instances))]
    return mix


def _mix_rate(mix, kws, tuning):
    with S.Engine(max_slots=len(mix)) as e:
        admit_specs(e, mix, tuning, kws=kws)
        return _day_rate(e)


def autotune(specs, chunks=CHUNK_CANDIDATES, tiles=TILE_CANDIDATES, mix_agents=300_000_000, log=None):
    """Coordinate descent on a mix of the set: the chunk size per size
    bucket, then the tile shape per size bucket (at most MAX_TILE_SHAPES
    distinct shapes counting the default one, which buckets outside the mix
    keep), largest share of agents first, keeping what gives the
    most agent-steps/s over a day."""
    mix = tuning_mix(specs, mix_agents)
    kws = [kw for _, kw in W.materialise_all(mix)]
    share = {}
    for sp in mix:
        share[bucket_of(sp.n0)] = share.get(bucket_of(sp.n0), 0) + sp.n0
    order = sorted(share, key=lambda x: -share[x])
    base = Tuning()
    if log:
        log(f"autotune: mix of {len(mix)} scenarios, {sum(share.values()) / 1e6:.0f}M agents")

    def rate_of(chunk, tile):
        return _mix_rate(mix, kws, Tuning(tile=base.tile, chunk=dict(chunk), tiles=dict(tile)))[0]

    chunk, _, t1 = coordinate_descent(order, chunks, {b: DEFAULT_CHUNK for b in order},
                                      lambda c: rate_of(c, {b: base.tile for b in order}), log=log, what="chunk")
    tile, best, t2 = coordinate_descent(order, [tuple(t) for t in tiles], {b: tuple(base.tile) for b in order},
                                        lambda t: rate_of(chunk, t),
                                        allowed=lambda t: len(set(t.values()) | {tuple(base.tile)}) <= MAX_TILE_SHAPES, log=log,
                                        what="tile")
    table = [{"round": "chunk", **r} for r in t1] + [{"round": "tile", **{k: (list(v) if isinstance(v, tuple) else v)
                                                                           for k, v in r.items()}} for r in t2]
    for r in table:  # JSON keys
        if "state" in r:
            r["state"] = {k: (list(v) if isinstance(v, tuple) else v) for k, v in r["state"].items()}
    return Tuning(tile=base.tile, chunk=chunk, tiles=tile, table=table)


def run_batched(specs, tuning, warm_days=1, timed_days=7, phys_factor=1.25, headroom=1.25, keep_rows=(), log=None):
    """All specs in one engine; times `timed_days` graph-launched days in one
    submission after `warm_days`, with no host synchronisation inside the
    timed region (instances that outgrow their buffers pause until the
    growth after it).  Returns measurements and the rows of the slots listed
    in keep_rows (spec indices)."""
    free0, total = S.device_memory()
    t_admit = time.perf_counter()
    with S.Engine(max_slots=len(specs)) as e:
        slots = admit_specs(e, specs, tuning, phys_factor)
        e.sync()
        admit_s = time.perf_counter() - t_admit
        free1, _ = S.device_memory()
        if log:
            log(f"batched: admitted {len(specs)} instances in {admit_s:.1f} s, "
                f"{(free0 - free1) / 2**30:.1f} GiB resident")
        for _ in range(warm_days):
            e.step_days(1)
            e.service(headroom)
        e.agent_steps(reset=True)
        t0 = time.perf_counter()
        e.step_days(timed_days)
        e.sync()
        dt = time.perf_counter() - t0
        steps = e.agent_steps()
        st, step, na = e.poll(len(specs))
        paused = int((st == S.GROW).sum())
        grown = e.service(headroom)
        # the fused pass alone over one more day, for the bandwidth figure:
        # 34 B per agent in and out of each iteration
        _, _, na0 = e.poll(len(specs))
        e.agent_steps(reset=True)
        wall_ms, pass_ms = e.time_pass(24)
        pass_steps = e.agent_steps()
        st, step, na = e.poll(len(specs))
        free2, _ = S.device_memory()
        rows = {i: e.rows(slots[i], 0, int(step[i]) + 1) for i in keep_rows}
    n_in = pass_steps
    n_out = pass_steps - int(na0.sum()) + int(na.sum())
    pass_bytes = 34.0 * (n_in + n_out)
    status = {int(k): int(v) for k, v in zip(*np.unique(st, return_counts=True))}
    return {"instances": len(specs), "agent_steps": steps, "seconds": dt, "agent_steps_per_s": steps / dt,
            "day_ms": dt * 1e3 / timed_days, "timed_days": timed_days, "paused_for_growth": paused,
            "grown": grown, "admit_s": admit_s, "resident_gib": (free0 - free2) / 2**30, "status": status,
            "agents_now": int(na.sum()), "pass_ms_per_iteration": pass_ms / 24, "pass_agent_steps": pass_steps,
            "pass_algorithmic_bytes": pass_bytes, "pass_GBs": pass_bytes / (pass_ms / 1e3) / 1e9}, rows


SMALL_RUN = 1_000_000   # below this many initial agents a lone run goes by day launches
SMALL_CHUNK = 512       # ... with chunks of this many agents
DAY_WORK_GROUPS = 256


def run_plain_engine(specs, days=1, warm_days=1, keep_rows=True, policy="default"):
    """Each spec alone in its own engine; agent-steps/s over the sample and
    each run's rows.  policy "default": the engine's default launch
    parameters (a pass and a prep per iteration, one day graph per day);
    "single-run": what bench.py uses for a lone run, day launches of 256
    work-groups and 512-agent chunks below SMALL_RUN initial agents."""
    import time
    steps = 0
    secs = 0.0
    rows = []
    for sp, kw in W.materialise_all(specs):
        with S.Engine(max_slots=1) as e:
            day = policy == "single-run" and sp.n0 < SMALL_RUN
            if day:
                e.set_chunk(SMALL_CHUNK, 0)
                e.set_day(DAY_WORK_GROUPS)
            slot = e.admit(**kw)
            for _ in range(warm_days):
                e.step_days(1)
            e.agent_steps(reset=True)
            if day:  # a day launch per day; the host clock around the timed days
                e.sync()
                t0 = time.perf_counter()
                e.step_days(days)
                e.sync()
                secs += time.perf_counter() - t0
            else:
                secs += e.time(24 * days, graph=True) / 1e3
            steps += e.agent_steps()
            if keep_rows:
                rows.append(e.rows(slot, 0, 24 * (warm_days + days) + 1))
    return {"instances": len(specs), "agent_steps": steps, "seconds": secs, "policy": policy,
            "agent_steps_per_s": steps / secs if secs else 0.0}, rows


def run_plain_icl(specs, days=1):
    """Each spec alone through the reference's driver loop; agent-steps/s
    over the sample and each run's rows."""
    steps = 0
    secs = 0.0
    rows = []
    for sp, kw in W.materialise_all(specs):
        r = S.run_abms(seed=kw["seed"], max_agents=sp.max_agents, steps=24 * days, temp=kw["temp"], env=kw["env"],
                       mode=S.MODE_CONFIG, dmr=kw["dmr"], layout=kw["layout"], agents=kw["agents"], export=False)
        steps += int(r["agent_steps"])
        secs += r["loop_ms"] / 1e3
        rows.append(r["rows"])
    return {"instances": len(specs), "agent_steps": steps, "seconds": secs,
            "agent_steps_per_s": steps / secs if secs else 0.0}, rows


def same_rows(a, b, n):
    """Rows 0..n-1 of two runs are identical (every printStats field)."""
    n = min(n, len(a), len(b))
    return bool(np.array_equal(a[:n], b[:n]))


def save_tuning(tuning, path):
    with open(path, "w") as f:
        json.dump(tuning.to_json(), f, indent=1)


def load_tuning(path):
    d = json.load(open(path))
    return Tuning(tile=tuple(d["tile"]), chunk={int(k): int(v) for k, v in d["chunk"].items()},
                  tiles={int(k): tuple(v) for k, v in d.get("tiles", {}).items()}, table=d.get("table", []))
