"""Full-size runs on the GPU, where the oracle would take minutes: the
reference defaults for a whole year (src/abms.c:15-19: 32,000 eggs,
CC = 160,000, 8,760 hourly iterations, T = 30 degC, blood-meal success 0.25,
capacity 6,000,000, include/abms.h:16).

The fused engine and the reference's driver loop over the icl boundary share
no kernel code (one fused pass on structure-of-arrays storage against the
reference's kernel sequence on its 24/16/8-byte records), so equal rows,
agents and histograms after 8,760 iterations pin each other at the size the
bit-exact oracle comparisons cannot reach; the oracle pins both on shorter
runs (test_gpu_parity.py).  The year is also checked against the paper's
emergent number: a population peak of about 1.4 million agents at these
settings (the SAMPO paper, Fig. 6; eggs are most of it).
"""
import numpy as np
import pytest

import sampo_amd as S

pytestmark = pytest.mark.gpu

STEPS, EGGS, CC, CAP, SEED = 8760, 32000, 160000.0, 6_000_000, 1


def test_reference_defaults_full_year_engine_equals_driver_loop():
    temp = np.full(STEPS, 30.0, np.float32)
    env = S.make_env(cc=CC)
    with S.Engine(max_slots=1) as e:
        slot = e.admit(mode=S.MODE_REFERENCE, seed=SEED, horizon=STEPS, temp=temp, env=env, n_eggs=EGGS,
                       max_agents=CAP)
        e.step_days(STEPS // 24)
        st, step, n = e.status(slot)
        rows_e = e.rows(slot, 0, STEPS)
        lay_e, ag_e = e.export(slot)
        hist_e = e.histogram(slot)
    assert st == S.DONE and step == STEPS
    g = S.run_abms(seed=SEED, n_eggs=EGGS, max_agents=CAP, steps=STEPS, temp=temp, env=env)
    assert g["status"] == 0 and g["steps_done"] == STEPS
    rows_g = g["rows"]
    assert np.array_equal(rows_e["count"], rows_g["count"])
    for f in ("l1de", "females", "potInfective", "numCycles", "sumCycles", "bites", "infectBites"):
        assert np.array_equal(rows_e[f], rows_g[f]), f
    assert np.array_equal(lay_e, g["layout"])
    idx = np.concatenate([np.arange(lay_e[r][0], lay_e[r][1]) for r in range(8)])
    for f in ("availableEggs", "cycleLength", "humanBloodmealCount", "delay", "numEggBatches", "aux"):
        assert np.array_equal(g["agents"][f][idx], ag_e[f]), f
    for f in ("ageInHours", "hoursInState", "isFemale", "csd"):
        assert np.array_equal(g["ages"][f][idx], ag_e[f]), f
    assert np.array_equal(hist_e, g["histogram"])
    # the paper's year at these settings: the population peaks near 1.4 million
    total = rows_e["count"].astype(np.int64).sum(axis=1)
    peak = int(total.max())
    print(f"population peak {peak} at iteration {int(total.argmax())}, agents at the end {n}", flush=True)
    assert 0.7e6 < peak < 2.8e6
