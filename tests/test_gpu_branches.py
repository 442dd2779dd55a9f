"""Every update branch on the GPU against closed forms (MI355X).

The hand-built cases of tests/branch_cases.py run through the fused engine
(one export per iteration) and through the reference driver loop over the
icl boundary (one run per prefix length); each path's fields, counts and
newborns must satisfy the closed-form checks computed from independent
random draws, and match the oracle bit for bit.  The age histogram of both
paths is checked against numpy.bincount over the exported adults.
"""
import numpy as np
import pytest

import sampo_amd as S
from oracle import pyoracle as O
from tests import branch_cases as B

pytestmark = pytest.mark.gpu


def species_of(case):
    sp = S.default_species()
    for k, v in case.species.items():
        sp[k] = v
    return sp


def run_engine(case):
    lay, ag = B.population(case.groups)
    with S.Engine(max_slots=1) as e:
        slot = e.admit(mode=S.MODE_CONFIG, seed=case.seed, horizon=case.steps + 1, temp=case.temp, env=case.env,
                       max_agents=1 << 16, dmr=(0.0, 0.0, 0.0, 0.0), layout=lay, agents=ag,
                       species=species_of(case))
        snaps = [e.export(slot)]
        for _ in range(case.steps):
            e.step(1)
            snaps.append(e.export(slot))
        rows = e.rows(slot, 0, case.steps + 1)
    return rows, snaps


def icl_dense(r):
    lay = r["layout"]
    idx = np.concatenate([np.arange(int(a), int(b), dtype=np.int64) for a, b in lay])
    out = np.zeros(idx.shape[0], dtype=B.AGENT)
    for f in ("availableEggs", "cycleLength", "humanBloodmealCount", "delay", "numEggBatches", "aux"):
        out[f] = r["agents"][f][idx]
    for f in ("ageInHours", "hoursInState", "isFemale", "csd"):
        out[f] = r["ages"][f][idx]
    return lay, out


def run_icl(case, tmp_path):
    lay, ag = B.population(case.groups)
    spf = None
    if case.species:
        spf = tmp_path / f"{case.name}.species"
        S.write_species(spf, species_of(case))
    snaps, rows = [(lay, ag)], None
    for t in range(1, case.steps + 2):
        r = S.run_abms(seed=case.seed, max_agents=1 << 16, steps=t, temp=case.temp, env=case.env,
                       mode=S.MODE_CONFIG, dmr=(0.0, 0.0, 0.0, 0.0), layout=lay, agents=ag, species=spf)
        assert r["steps_done"] == t
        if t <= case.steps:
            snaps.append(icl_dense(r))
        rows = r["rows"]
    return rows, snaps


def run_oracle(case):
    from tests.test_oracle_branches import run_oracle as ro
    return ro(case)


STATS = ("l1de", "females", "potInfective", "numCycles", "sumCycles", "bites", "infectBites")


def same(a, b):
    (ra, sa), (rb, sb) = a, b
    assert len(ra) == len(rb) and np.array_equal(ra["count"], rb["count"])
    for f in STATS:
        assert np.array_equal(ra[f], rb[f]), f
    for (la, xa), (lb, xb) in zip(sa, sb):
        assert np.array_equal(la, lb)
        for f in B.AGENT.names:
            assert np.array_equal(xa[f], xb[f]), f


@pytest.mark.parametrize("make", B.CASES, ids=lambda f: f.__name__[5:])
def test_engine_update_branch_closed_form(make):
    case = make()
    print(f"engine branch case {case.name} seed {case.seed}", flush=True)
    got = run_engine(case)
    case.check(*got)
    same(got, run_oracle(case))


@pytest.mark.parametrize("make", B.CASES, ids=lambda f: f.__name__[5:])
def test_icl_update_branch_closed_form(make, tmp_path):
    case = make()
    print(f"icl branch case {case.name} seed {case.seed}", flush=True)
    got = run_icl(case, tmp_path)
    case.check(*got)
    same(got, run_oracle(case))


def test_histograms_are_bincount_of_adult_days():
    lay, ag = B.population(B.histogram_population())
    temp, env = B.steady_temp(2), B.env()
    with S.Engine(max_slots=1) as e:
        slot = e.admit(mode=S.MODE_CONFIG, seed=3, horizon=2, temp=temp, env=env, max_agents=1 << 16,
                       dmr=(0.0, 0.0, 0.0, 0.0), layout=lay, agents=ag)
        e.step(1)
        l1, a1 = e.export(slot)
        h = e.histogram(slot)
    want = B.histogram_of(l1, a1)
    assert want[99] >= 4 and np.array_equal(h, want)
    r = S.run_abms(seed=3, max_agents=1 << 16, steps=1, temp=temp, env=env, mode=S.MODE_CONFIG,
                   dmr=(0.0, 0.0, 0.0, 0.0), layout=lay, agents=ag)
    assert np.array_equal(r["histogram"], B.histogram_of(*icl_dense(r)))
