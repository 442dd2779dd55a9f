"""The batched and serving workloads at their real sizes (MI355X).

  ensemble  all 1,024 replicates (initial populations uniform 1e5-1e6) in one
            engine for one day;
  sweep     all 4,096 scenarios of the uneven parameter sweep (initial
            population log-uniform 1e3-4e6, per-scenario temperature and
            mortality) in one engine for one day;
  serving   a one-second window of the Poisson(50/s) arrival process with
            its size and horizon ranges (1e4-1e7 agents, 30-365 days), served
            to completion.
Each instance must be bit-identical to its run alone: a sample of the
batched instances (including the smallest and the largest) is re-run in an
engine of its own, and the two smallest are also run through the oracle.
"""
import numpy as np
import pytest

import sampo_amd as S
from oracle import pyoracle as O
from sampo_amd import batch as B
from sampo_amd import serve as V
from sampo_amd import workloads as W

pytestmark = pytest.mark.gpu

STATS = ("l1de", "females", "potInfective", "numCycles", "sumCycles", "bites", "infectBites")


def same_rows(a, b, n):
    if len(a) < n or len(b) < n:
        return False
    return np.array_equal(a["count"][:n], b["count"][:n]) and all(np.array_equal(a[f][:n], b[f][:n]) for f in STATS)


def sample_of(specs, k=8):
    by_size = sorted(range(len(specs)), key=lambda i: specs[i].n0)
    spread = np.linspace(0, len(specs) - 1, k - 2).round().astype(int).tolist()
    return sorted(set([by_size[0], by_size[-1]] + spread))


def batched_day(specs, keep):
    """Every spec in one engine, one graph-launched day; the rows of keep."""
    tun = B.Tuning()
    with S.Engine(max_slots=len(specs)) as e:
        slots = B.admit_specs(e, specs, tun, phys_factor=1.25)
        e.step_days(1)
        e.service(1.25)
        st, step, na = e.poll(len(specs))
        rows = {i: e.rows(slots[i], 0, 25) for i in keep}
    return st, step, na, rows


def solo_rows(spec):
    kw = W.materialise(spec)
    with S.Engine(max_slots=1) as e:
        slot = e.admit(**kw)
        e.step_days(1)
        return e.rows(slot, 0, 25)


def oracle_rows(spec, iterations=24):
    kw = W.materialise(spec)
    sim = O.Sim(temp=kw["temp"], env=kw["env"], capacity=spec.max_agents, seed=spec.seed, layout=kw["layout"],
                agents=kw["agents"], mode=1, dmr=spec.dmr, threads=16)
    return sim.run(iterations)


def check_set(specs, what):
    keep = sample_of(specs)
    by_size = sorted(range(len(specs)), key=lambda i: specs[i].n0)
    print(f"{what}: {len(specs)} instances, {sum(s.n0 for s in specs) / 1e9:.3f}e9 initial agents", flush=True)
    st, step, na, rows = batched_day(specs, keep + by_size[:2])
    assert (st == S.RUNNING).all() and (step == 24).all(), f"{what}: an instance stopped"
    for i in keep:
        print(f"{what}: instance {i} (n0 {specs[i].n0}) alone", flush=True)
        assert same_rows(rows[i], solo_rows(specs[i]), 25), f"{what}: instance {i} differs from its run alone"
    for i in by_size[:2]:
        print(f"{what}: instance {i} (n0 {specs[i].n0}) against the oracle", flush=True)
        assert same_rows(rows[i], oracle_rows(specs[i]), 24), f"{what}: instance {i} differs from the oracle"


def test_ensemble_full_set_is_identical_to_solo_runs():
    check_set(W.ensemble(count=1024), "ensemble")


def test_sweep_full_set_is_identical_to_solo_runs():
    check_set(W.sweep(count=4096), "sweep")


def test_serving_window_each_request_same_alone_and_in_the_mix():
    specs = W.serving(rate=50.0, duration=1.0)
    assert len(specs) >= 20
    n0 = np.array([s.n0 for s in specs])
    assert n0.min() >= 1e4 and n0.max() <= 1e7
    print(f"serving: {len(specs)} requests, n0 {n0.min()}..{n0.max()}, "
          f"horizons {min(s.days for s in specs)}..{max(s.days for s in specs)} days", flush=True)
    srv = V.Server(max_slots=256, keep_rows=True)
    results, summary = srv.run(specs, clock="wall")
    print(f"serving: p50 {summary.get('p50')} p99 {summary.get('p99')} s, "
          f"{summary['agent_steps_per_s']:.3e} agent-steps/s", flush=True)
    assert summary["completed"] == len(specs)
    for sp, r in zip(specs, results):
        # a run ends at its horizon, or earlier where the reference stops: more
        # than 0.95 x capacity agents or none (src/abms.c:305-308)
        assert (r.status == S.DONE and r.steps == sp.horizon) or \
            (r.status in (S.CAPACITY, S.EXTINCT) and r.steps < sp.horizon), (sp.index, r.status, r.steps)
    order = sorted(range(len(specs)), key=lambda i: specs[i].n0)
    stopped = [i for i, r in enumerate(results) if r.status != S.DONE]
    pick = sorted(set([order[0], order[-1], max(range(len(specs)), key=lambda i: specs[i].days)] + stopped[:2] +
                      list(range(0, len(specs), max(1, len(specs) // 3)))))
    for i in pick:
        print(f"serving: request {i} (n0 {specs[i].n0}, {specs[i].days} days) alone", flush=True)
        alone = V.run_alone(specs[i])
        assert np.array_equal(alone, results[i].rows), f"request {i} differs from its run alone"
    i = order[0]
    n = results[i].steps
    print(f"serving: request {i} against the oracle, {n} iterations", flush=True)
    assert same_rows(results[i].rows, oracle_rows(specs[i], n), n)
