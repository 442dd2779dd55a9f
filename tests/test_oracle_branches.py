"""The oracle's update branches against closed forms (CPU).

Every branch of kernel/updateAgents.cl, the egg creation of
kernel/initAgents.cl and the laying/egg-batch functions of
include/properties.h run on hand-built agents (tests/branch_cases.py) whose
expected fields, counts and newborns are recomputed from independent random
draws and the reference's constants.  The age histogram is checked against
numpy.bincount over the exported adults.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import branch_cases as B


def run_oracle(case):
    lay, ag = B.population(case.groups)
    sp = O.default_species()
    for k, v in case.species.items():
        sp[k] = v
    sim = O.Sim(temp=case.temp, env=case.env, capacity=1 << 16, seed=case.seed, layout=lay, agents=ag, mode=1,
                dmr=(0.0, 0.0, 0.0, 0.0), species=sp)
    rows, snaps = [], [(lay, ag)]
    for _ in range(case.steps):
        rows.append(sim.run(1)[0])
        l, a, _ = sim.export()
        snaps.append((l, a))
    rows.append(sim.run(1)[0])  # the row at the start of iteration `steps`
    return np.array(rows, dtype=O.ROW), snaps


@pytest.mark.parametrize("make", B.CASES, ids=lambda f: f.__name__[5:])
def test_update_branch_closed_form(make):
    case = make()
    rows, snaps = run_oracle(case)
    case.check(rows, snaps)


def test_histogram_is_bincount_of_adult_days():
    lay, ag = B.population(B.histogram_population())
    sim = O.Sim(temp=B.steady_temp(2), env=B.env(), capacity=1 << 16, seed=3, layout=lay, agents=ag, mode=1,
                dmr=(0.0, 0.0, 0.0, 0.0))
    sim.run(1)
    l1, a1, _ = sim.export()
    want = B.histogram_of(l1, a1)
    assert np.array_equal(sim.histogram(), want)
    assert want[99] >= 4 and want[0] >= 1 and want[1] >= 2  # the cap and both sides of a day edge
    # and over a whole configuration-mode population
    import sampo_amd as S
    temp = S.gen_temperature(5, 4)
    lay, ag = S.gen_population(5, 4000, float(temp[0]))
    sim = O.Sim(temp=temp, env=S.make_env(cc=20000.0), capacity=200000, seed=5, layout=lay, agents=ag, mode=1,
                dmr=(0.05, 0.03, 0.02, 0.1))
    sim.run(24 * 4)
    l2, a2, _ = sim.export()
    assert np.array_equal(sim.histogram(), B.histogram_of(l2, a2))
