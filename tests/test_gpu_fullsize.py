"""Full-size runs on the GPU against oracle fixtures.

tests/golden/make_fullsize.py ran the oracle (the C restatement of the
reference, on the CPU of the build container) at three sizes and kept every
printStats row, the final layout, a SHA-256 of the final compacted agent
array, per-field sums and the age histogram:
  R   the reference defaults for a whole year (src/abms.c:13-19: 32,000 eggs,
      CC = 160,000, 8,760 hourly iterations, T = 30 degC, blood-meal success
      0.25, capacity 6,000,000, include/abms.h:16);
  C0  the small single run, 100,000 initial agents, 365 days;
  C1  the large single run, 32,000,000 initial agents in a 128,000,000
      capacity, its first 2 days.
The fused engine (and, for R and C0, the reference's driver loop over the icl
boundary) must reproduce them bit for bit.  Both GPU paths share the model
and numeric headers (sampo_model.h, sampo_canon.h) with each other and with
the oracle, so these fixtures -- written by the oracle, whose functions are
pinned independently (test_oracle_*.py) -- are what pins them at full size.
C1 is also checked by its invariants at full size: count conservation,
64-aligned range starts and the calcStats sums by brute force.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import sampo_amd as S
from sampo_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    z = np.load(os.path.join(GOLDEN, f"fullsize_{name}.npz"))
    return {k: z[k] for k in z.files}, json.loads(str(z["meta"]))


def digest(layout, agents):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(layout, dtype="<u4").tobytes())
    h.update(np.ascontiguousarray(agents).tobytes())
    return h.hexdigest()


STATS = ("l1de", "females", "potInfective", "numCycles", "sumCycles", "bites", "infectBites")


def rows_u32(rows):
    """The printStats fields of rows (engine, driver loop or oracle records)
    as an (n, 15) array."""
    return np.column_stack([rows["count"]] + [rows[f] for f in STATS]).astype(np.uint32)


def check_against(fx, rows, layout, agents, hist, what):
    want = fx["rows"]
    got = rows_u32(rows)
    assert got.shape == want.shape, what
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, f"{what}: first differing row {bad[0] if len(bad) else None}"
    assert np.array_equal(layout, fx["layout"]), what
    words = np.ascontiguousarray(agents).view(np.uint32).reshape(len(agents), -1).astype(np.uint64)
    sums = words.sum(axis=0)
    assert np.array_equal(sums, fx["field_sums"]), f"{what}: per-field sums {sums} vs {fx['field_sums']}"
    assert digest(layout, agents) == str(fx["sha256"]), what
    assert np.array_equal(hist, fx["histogram"]), what


def icl_dense(g):
    lay = g["layout"]
    idx = np.concatenate([np.arange(lay[r][0], lay[r][1]) for r in range(8)])
    out = np.zeros(idx.shape[0], dtype=S.AGENT_DTYPE)
    for f in ("availableEggs", "cycleLength", "humanBloodmealCount", "delay", "numEggBatches", "aux"):
        out[f] = g["agents"][f][idx]
    for f in ("ageInHours", "hoursInState", "isFemale", "csd"):
        out[f] = g["ages"][f][idx]
    return out


def test_reference_defaults_full_year_engine_and_driver_loop_equal_the_oracle():
    fx, m = fixture("R")
    steps = m["iterations"]
    temp = np.full(steps, m["temp"], np.float32)
    env = S.make_env(cc=m["cc"])
    print("R: engine, full year", flush=True)
    with S.Engine(max_slots=1) as e:
        slot = e.admit(mode=S.MODE_REFERENCE, seed=m["seed"], horizon=steps, temp=temp, env=env,
                       n_eggs=m["n_eggs"], max_agents=m["capacity"])
        e.step_days(steps // 24)
        st, step, n = e.status(slot)
        rows = e.rows(slot, 0, steps)
        lay, ag = e.export(slot)
        hist = e.histogram(slot)
    assert st == S.DONE and step == steps
    check_against(fx, rows, lay, ag, hist, "engine")
    print("R: icl driver loop, full year", flush=True)
    g = S.run_abms(seed=m["seed"], n_eggs=m["n_eggs"], max_agents=m["capacity"], steps=steps, temp=temp, env=env)
    assert g["status"] == 0 and g["steps_done"] == steps
    check_against(fx, g["rows"], g["layout"], icl_dense(g), g["histogram"], "icl driver")
    # the paper's year at these settings (Fig. 6): the population peaks near 1.4 million
    total = rows_u32(rows)[:, :8].astype(np.int64).sum(axis=1)
    assert 1.0e6 < int(total.max()) < 2.0e6


def test_small_single_run_full_year_equals_the_oracle():
    fx, m = fixture("C0")
    w = W.SINGLE_RUN["small"]
    temp, env, layout, agents = W.single_run_inputs(w)
    steps = m["iterations"]
    print("C0: engine, 100k agents, full year", flush=True)
    with S.Engine(max_slots=1) as e:
        e.set_chunk(512, 0)
        slot = e.admit(mode=S.MODE_CONFIG, seed=w["seed"], horizon=steps, temp=temp, env=env,
                       max_agents=w["max_agents"], dmr=W.DMR, layout=layout, agents=agents)
        e.step_days(steps // 24)
        st, step, _ = e.status(slot)
        rows = e.rows(slot, 0, steps)
        lay, ag = e.export(slot)
        hist = e.histogram(slot)
    assert st == S.DONE and step == steps
    check_against(fx, rows, lay, ag, hist, "engine")
    print("C0: icl driver loop, full year", flush=True)
    g = S.run_abms(seed=w["seed"], max_agents=w["max_agents"], steps=steps, temp=temp, env=env, mode=S.MODE_CONFIG,
                   dmr=W.DMR, layout=layout, agents=agents)
    assert g["status"] == 0 and g["steps_done"] == steps
    check_against(fx, g["rows"], g["layout"], icl_dense(g), g["histogram"], "icl driver")


def test_large_single_run_first_days_equal_the_oracle():
    fx, m = fixture("C1")
    w = W.SINGLE_RUN["large"]
    temp, env, layout, agents = W.single_run_inputs(w)
    steps = m["iterations"]
    print(f"C1: engine, {w['n0']} agents in {w['max_agents']}, {steps} iterations", flush=True)
    with S.Engine(max_slots=1) as e:
        slot = e.admit(mode=S.MODE_CONFIG, seed=w["seed"], horizon=24 * w["days"], temp=temp, env=env,
                       max_agents=w["max_agents"], dmr=W.DMR, layout=layout, agents=agents)
        del agents
        e.step_days(steps // 24)
        st, step, n = e.status(slot)
        rows = e.rows(slot, 0, steps + 1)
        lay, ag = e.export(slot)
        hist = e.histogram(slot)
    assert st == S.RUNNING and step == steps
    check_against(fx, rows[:steps], lay, ag, hist, "engine")
    # invariants at full size, from the exported population after the last iteration
    cnt = (lay[:, 1] - lay[:, 0]).astype(np.int64)
    assert cnt.sum() == n == len(ag) and np.array_equal(rows[steps]["count"], cnt)
    assert lay[0][0] == 0 and all(lay[s][0] == ((lay[s - 1][1] + 64) // 64) * 64 for s in range(1, 8))
    k_l0, k_l1 = cnt[0], cnt[0] + cnt[1]
    l1de = np.floor(ag["ageInHours"][k_l0:k_l1] / np.float32(24.0)).astype(np.uint64).sum()
    adult = ag[k_l1 + cnt[2]:]
    fem = adult["isFemale"] == 1
    assert rows[steps]["l1de"] == l1de
    assert rows[steps]["females"] == fem.sum()
    assert rows[steps]["potInfective"] == (fem & (adult["csd"] >= 1.0)).sum()
