"""Whole-iteration checks of the oracle (CPU).

The reference has two interchangeable scan paths (src/abms.c:404-439): the
GPU path (preScan + Bolt scan-by-key) and the CPU-device path (eight
serialScan calls).  Their restatements share nothing but the agent kernels,
so identical trajectories pin the scan/compaction restatement; the
invariants below pin layout, reductions and update rules.
"""
import numpy as np
import pytest

import sampo_amd as S
from oracle import pyoracle as O

FIELDS = O.AGENT.names


def ref_run(n_eggs, steps, cc, scan_path, seed=1, temp=30.0):
    t = np.full(steps, temp, np.float32)
    sim = O.Sim(temp=t, env=S.make_env(cc=cc), capacity=200000, seed=seed, n_eggs=n_eggs, scan_path=scan_path)
    rows = sim.run(steps)
    return sim, rows


def config_sim(n0, steps, seed, scan_path=0, cc=None, dmr=(0.05, 0.03, 0.02, 0.1)):
    temp = S.gen_temperature(seed, (steps + 23) // 24)[:steps]
    lay, ag = S.gen_population(seed, n0, float(temp[0]))
    cc = cc if cc is not None else 5.0 * n0
    sim = O.Sim(temp=temp, env=S.make_env(cc=cc), capacity=40 * n0 + 4096, seed=seed, layout=lay, agents=ag,
                mode=1, dmr=dmr, scan_path=scan_path)
    return sim


def test_bolt_and_serial_scan_paths_agree_reference_mode():
    a, ra = ref_run(1500, 24 * 30, 8000.0, 0)
    b, rb = ref_run(1500, 24 * 30, 8000.0, 1)
    assert np.array_equal(ra, rb)
    la, xa, sa = a.export()
    lb, xb, sb = b.export()
    assert np.array_equal(la, lb) and np.array_equal(xa, xb) and np.array_equal(sa, sb)


def test_bolt_and_serial_scan_paths_agree_config_mode():
    a = config_sim(3000, 24 * 20, 7, scan_path=0)
    b = config_sim(3000, 24 * 20, 7, scan_path=1)
    assert np.array_equal(a.run(24 * 20), b.run(24 * 20))
    assert np.array_equal(a.export()[1], b.export()[1])


def check_invariants(sim, rows):
    lay, ag, st = sim.export()
    # ranges: egg range starts at 0; later starts are the next multiple of 64 after the previous end
    assert lay[0][0] == 0
    for s in range(1, 8):
        assert lay[s][0] == ((lay[s - 1][1] + 64) // 64) * 64
    # every agent sits in the range of its state
    k = 0
    for s in range(8):
        n = lay[s][1] - lay[s][0]
        assert np.all(st[k:k + n] == (1 << s))
        k += n
    # hours stay integral; males never leave mate seeking
    assert np.all(ag["ageInHours"] == np.floor(ag["ageInHours"]))
    assert np.all(ag["hoursInState"] == np.floor(ag["hoursInState"]))
    adult_from_bms = int(lay[5][0] - lay[0][0])
    start_bms = int(sum(lay[s][1] - lay[s][0] for s in range(5)))
    assert np.all(ag["isFemale"][start_bms:] == 1)
    # the histogram counts every adult once
    n_adult = int(sum(lay[s][1] - lay[s][0] for s in range(3, 8)))
    assert int(sim.histogram().sum()) == n_adult
    return lay, ag, st


def test_reference_mode_invariants_and_reductions():
    sim, rows = ref_run(2000, 24 * 25, 10000.0, 0)
    lay, ag, st = check_invariants(sim, rows)
    # the next row's reductions equal brute-force sums over the exported agents
    nxt = sim.run(1)[0]
    counts = lay[:, 1] - lay[:, 0]
    assert np.array_equal(nxt["count"], counts)
    k1 = int(counts[0])
    larvae = ag[k1:k1 + int(counts[1])]
    assert nxt["l1de"] == int(np.sum(np.floor(larvae["ageInHours"] / np.float32(24)).astype(np.uint64)))
    adults = ag[int(counts[:3].sum()):]
    fem = adults["isFemale"] == 1
    assert nxt["females"] == int(fem.sum())
    assert nxt["potInfective"] == int((fem & (adults["csd"] >= 1.0)).sum())
    # conservation of counted agents is implied by the row counts: no agent appears twice
    assert len(ag) == int(counts.sum())


def test_population_dynamics_plausible():
    """Paper Fig. 5: eggs per batch ~ N(170, 30) * 0.8^k; the reference model
    grows from eggs to a stationary population with adults and bites."""
    sim, rows = ref_run(2000, 24 * 60, 10000.0, 0)
    last = rows[-1]
    assert last["count"][3:].sum() > 0 and last["bites"] >= 0
    lay, ag, st = sim.export()
    grav = ag[int((lay[:7, 1] - lay[:7, 0]).sum()):]
    first_batch = grav[(grav["numEggBatches"] == 1) & (grav["aux"] == 0)]["availableEggs"]
    if len(first_batch) > 20:
        assert 110 < first_batch.mean() < 230


def test_capacity_and_extinction_statuses():
    sim, rows = ref_run(2000, 24 * 40, 10000.0, 0)
    # a tiny capacity trips the 95% check (src/abms.c:305)
    t = np.full(400, 30.0, np.float32)
    small = O.Sim(temp=t, env=S.make_env(cc=10000.0), capacity=2100, seed=1, n_eggs=2000)
    small.run(400)
    assert small.status == 1
    # daily mortality of 1 for every stage empties the population (src/abms.c:307)
    dead = config_sim(500, 48, 3, dmr=(1.0, 1.0, 1.0, 1.0))
    r = dead.run(48)
    assert dead.status == 2 and len(r) < 48
