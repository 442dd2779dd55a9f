"""Write the full-size oracle fixtures tests/golden/fullsize_<name>.npz.

Each fixture is what the oracle (the C restatement of the reference, run on
the CPU here) produces for one workload at its real size:
  rows       every printStats row of the run (src/abms.c:67-81 fields),
  layout     the population ranges after the last iteration,
  sha256     of the layout and the compacted agent array after the last
             iteration (dense, reference index order, the 40-byte records
             of include/sampo_engine.h sampo_agent),
  field_sums per-field sums of that array's 32-bit words (to say which field
             differs when the hash does),
  histogram  ageHistogram after the last iteration,
  meta       the workload's parameters (JSON).

Workloads:
  R   the reference defaults (src/abms.c:13-19, include/abms.h:16): 32,000
      initial eggs, carrying capacity 160,000, 30 degC, the reference
      environment, capacity 6,000,000, seed 1, 8,760 iterations;
  C0  the small single run (sampo_amd.workloads.SINGLE_RUN["small"]):
      100,000 initial agents, 365 days;
  C1  the large single run (SINGLE_RUN["large"]), 32,000,000 initial agents,
      capacity 128,000,000: its first 2 days.

Run in the build container:
    python tests/golden/make_fullsize.py [R C0 C1]
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import sampo_amd as S  # noqa: E402
from oracle import pyoracle as O  # noqa: E402
from sampo_amd import workloads as W  # noqa: E402

R = dict(n_eggs=32_000, cc=160_000.0, temp=30.0, capacity=6_000_000, seed=1, iterations=8760)


def digest(layout, agents):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(layout, dtype="<u4").tobytes())
    h.update(np.ascontiguousarray(agents).tobytes())
    return h.hexdigest()


def field_sums(agents):
    words = np.ascontiguousarray(agents).view(np.uint32).reshape(len(agents), -1).astype(np.uint64)
    return words.sum(axis=0)


def save(name, rows, sim, meta):
    lay, ag, _ = sim.export()
    out = os.path.join(HERE, f"fullsize_{name}.npz")
    stats = ("l1de", "females", "potInfective", "numCycles", "sumCycles", "bites", "infectBites")
    table = np.column_stack([rows["count"]] + [rows[f] for f in stats]).astype(np.uint32)
    np.savez_compressed(out, rows=table, layout=lay,
                        sha256=np.array(digest(lay, ag)), field_sums=field_sums(ag), histogram=sim.histogram(),
                        meta=np.array(json.dumps(meta)))
    print(f"{out}: {len(rows)} rows, {len(ag)} agents at the end", flush=True)


def make_r(threads):
    it = R["iterations"]
    temp = np.full(it, R["temp"], np.float32)
    sim = O.Sim(temp=temp, env=S.make_env(cc=R["cc"]), capacity=R["capacity"], seed=R["seed"], n_eggs=R["n_eggs"],
                threads=threads)
    rows = sim.run(it)
    assert len(rows) == it, "reference run stopped early"
    save("R", rows, sim, dict(mode="reference", **R))


def make_config(name, key, days, threads):
    w = W.SINGLE_RUN[key]
    temp, env, layout, agents = W.single_run_inputs(w)
    n = 24 * days
    sim = O.Sim(temp=temp, env=env, capacity=w["max_agents"], seed=w["seed"], layout=layout, agents=agents, mode=1,
                dmr=W.DMR, threads=threads)
    del agents
    rows = sim.run(n)
    assert len(rows) == n, f"{name} stopped early"
    meta = {k: v for k, v in w.items() if k != "desc"}
    meta.update(mode="config", workload=key, iterations=n, dmr=list(W.DMR))
    save(name, rows, sim, meta)


def main(which):
    threads = os.cpu_count() or 1
    for name in which:
        t0 = time.perf_counter()
        if name == "R":
            make_r(threads)
        elif name == "C0":
            make_config("C0", "small", 365, threads)
        elif name == "C1":
            make_config("C1", "large", 2, threads)
        else:
            raise SystemExit(f"unknown fixture {name}")
        print(f"{name}: {time.perf_counter() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["R", "C0", "C1"])
