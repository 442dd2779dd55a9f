"""Day launches (sampo_engine_set_day): a lone instance advanced by one
cooperative launch per simulated day instead of two launches per iteration.

Bar: bit-exact.  Every printStats row, the final ranges, every agent field
in reference index order and the histogram equal the oracle's (and the
per-iteration launches' where the oracle would be slow), for launches of 1
to 256 work-groups, partial days, a horizon that ends inside a launch,
capacity and extinction, growth of the instance's buffers between launches,
interventions, the full-year small run, and an instance that shares the
engine with another one (then the per-iteration launches advance both).
"""
import numpy as np
import pytest

import sampo_amd as S
from oracle import pyoracle as O
from test_gpu_fullsize import check_against, fixture
from test_gpu_parity import agents_equal, config_inputs, intervention_env, ref_inputs, rows_equal
from sampo_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("groups", [1, 5, 64, 256])
def test_day_reference_mode_matches_oracle(groups):
    days, n_eggs, cc, seed = 6, 3200, 16000.0, 1
    steps = 24 * days
    temp, env = ref_inputs(steps, cc)
    sim = O.Sim(temp=temp, env=env, capacity=400_000, seed=seed, n_eggs=n_eggs)
    orows = sim.run(steps)
    print(f"day launches of {groups} work-groups, reference mode", flush=True)
    with S.Engine(max_slots=1) as e:
        e.set_day(groups)
        slot = e.admit(mode=S.MODE_REFERENCE, seed=seed, horizon=steps, temp=temp, env=env, n_eggs=n_eggs,
                       max_agents=400_000)
        e.step_days(days)
        st, step, _ = e.status(slot)
        assert st == S.DONE and step == steps
        rows_equal(orows, e.rows(slot, 0, steps))
        lay_o, ag_o, _ = sim.export()
        lay_g, ag_g = e.export(slot)
        assert np.array_equal(lay_o, lay_g)
        agents_equal(ag_o, ag_g)
        assert np.array_equal(sim.histogram(), e.histogram(slot))


def test_day_partial_launches_and_interventions_match_oracle():
    """Configuration mode with every intervention, advanced by launches of
    7, 24, 24, 1 and 40 iterations (step() splits into launches of <= 24)."""
    days, seed, n0 = 4, 23, 8000
    steps = 24 * days
    temp, lay, ag = config_inputs(n0, days, seed)
    env = intervention_env(steps, 0.025 * n0, seed)
    dmr = (0.05, 0.03, 0.02, 0.1)
    sim = O.Sim(temp=temp, env=env, capacity=400_000, seed=seed, layout=lay, agents=ag, mode=1, dmr=dmr)
    orows = sim.run(steps)
    print("day launches, configuration mode with interventions", flush=True)
    with S.Engine(max_slots=1) as e:
        e.set_day(96)
        slot = e.admit(mode=S.MODE_CONFIG, seed=seed, horizon=steps, temp=temp, env=env, max_agents=400_000,
                       dmr=dmr, layout=lay, agents=ag)
        e.step(7)
        e.step_days(2)
        e.step(1)
        e.step(40)
        st, step, _ = e.status(slot)
        assert st == S.DONE and step == steps
        rows_equal(orows, e.rows(slot, 0, steps))
        lay_o, ag_o, _ = sim.export()
        lay_g, ag_g = e.export(slot)
        assert np.array_equal(lay_o, lay_g)
        agents_equal(ag_o, ag_g)


def plain_and_day(admit, n_iter, groups, stepper):
    """(rows, status, step, layout, agents) of the same admission advanced by
    per-iteration launches and by day launches."""
    out = []
    for g in (0, groups):
        with S.Engine(max_slots=1) as e:
            e.set_day(g)
            slot = e.admit(**admit)
            stepper(e, slot)
            st, step, _ = e.status(slot)
            out.append((e.rows(slot, 0, n_iter + 1), st, step, *e.export(slot)))
    return out


def test_day_horizon_capacity_and_extinction_match_plain_launches():
    temp, env = ref_inputs(30, 20000.0)
    # a horizon that ends inside a launch: the last launch stops at it
    admit = dict(mode=S.MODE_REFERENCE, seed=4, horizon=30, temp=temp, env=env, n_eggs=2000, max_agents=100_000)
    a, b = plain_and_day(admit, 30, 32, lambda e, s: e.step_days(3))
    assert a[1] == b[1] == S.DONE and a[2] == b[2] == 30
    rows_equal(a[0], b[0])
    assert np.array_equal(a[3], b[3])
    agents_equal(a[4], b[4])
    # capacity: the launch ends at the iteration whose check fails
    steps = 24 * 30
    temp, env = ref_inputs(steps, 20000.0)
    admit = dict(mode=S.MODE_REFERENCE, seed=1, horizon=steps, temp=temp, env=env, n_eggs=2000, max_agents=2200)
    a, b = plain_and_day(admit, steps, 16, lambda e, s: e.step_days(30))
    assert a[1] == b[1] == S.CAPACITY and a[2] == b[2] < steps
    rows_equal(a[0], b[0])
    # extinction
    t2, lay, ag = config_inputs(400, 2, 5)
    admit = dict(mode=S.MODE_CONFIG, seed=5, horizon=48, temp=t2, env=S.make_env(cc=2000.0), max_agents=10000,
                 dmr=(1.0, 1.0, 1.0, 1.0), layout=lay, agents=ag)
    a, b = plain_and_day(admit, 48, 8, lambda e, s: e.step_days(2))
    assert a[1] == b[1] == S.EXTINCT and a[2] == b[2] < 48
    rows_equal(a[0], b[0])


def test_day_growth_between_launches_matches_oracle():
    """An instance admitted with only as many slots as its first population
    pauses inside a launch when its output would not fit, grows, resumes."""
    days, n0, seed = 30, 4000, 40
    steps = 24 * days
    temp, lay, ag = config_inputs(n0, days, seed)
    env = S.make_env(cc=0.025 * n0)
    dmr = (0.05, 0.03, 0.02, 0.1)
    sim = O.Sim(temp=temp, env=env, capacity=40 * n0, seed=seed, layout=lay, agents=ag, mode=1, dmr=dmr)
    orows = sim.run(steps)
    print("day launches with growth", flush=True)
    with S.Engine(max_slots=1) as e:
        e.set_day(64)
        slot = e.admit(mode=S.MODE_CONFIG, seed=seed, horizon=steps, temp=temp, env=env, max_agents=40 * n0,
                       phys_slots=n0, dmr=dmr, layout=lay, agents=ag)
        grown, paused, rounds = 0, False, 0
        while True:
            st, step, _ = e.status(slot)
            if st == S.DONE:
                break
            assert st in (S.RUNNING, S.GROW), st
            paused |= st == S.GROW
            grown += e.service(1.0)
            e.step_days(1)
            rounds += 1
            assert rounds < 4 * days
        assert grown > 0 and paused
        rows_equal(orows, e.rows(slot, 0, steps))
        lay_o, ag_o, _ = sim.export()
        lay_g, ag_g = e.export(slot)
        assert np.array_equal(lay_o, lay_g)
        agents_equal(ag_o, ag_g)


def test_day_instance_sharing_the_engine_matches_oracle():
    """A day-launch instance with a second instance beside it: the
    per-iteration launches advance both (its even cut included), and after
    the other is released day launches take over again."""
    days, seed = 4, 11
    steps = 24 * days
    temp, env = ref_inputs(steps, 12000.0)
    sims = [O.Sim(temp=temp, env=env, capacity=400_000, seed=seed + i, n_eggs=1500 + 700 * i) for i in range(2)]
    orows = [s.run(steps) for s in sims]
    with S.Engine(max_slots=2) as e:
        e.set_day(40)
        a = e.admit(mode=S.MODE_REFERENCE, seed=seed, horizon=steps, temp=temp, env=env, n_eggs=1500,
                    max_agents=400_000)
        e.step_days(1)
        b = e.admit(mode=S.MODE_REFERENCE, seed=seed + 1, horizon=steps, temp=temp, env=env, n_eggs=2200,
                    max_agents=400_000)
        # b starts one day behind: advance it alone for a day
        e.hold([True, False])
        e.step_days(1)
        e.hold([False, False])
        e.step_days(2)
        rb = e.rows(b, 0, 3 * 24)
        e.release(b)
        e.step_days(1)
        st, step, _ = e.status(a)
        assert st == S.DONE and step == steps
        rows_equal(orows[0], e.rows(a, 0, steps))
        rows_equal(orows[1][:3 * 24], rb)
        lay_o, ag_o, _ = sims[0].export()
        lay_g, ag_g = e.export(a)
        assert np.array_equal(lay_o, lay_g)
        agents_equal(ag_o, ag_g)


def test_day_small_single_run_full_year_equals_the_oracle():
    fx, m = fixture("C0")
    w = W.SINGLE_RUN["small"]
    temp, env, layout, agents = W.single_run_inputs(w)
    steps = m["iterations"]
    print("C0 by day launches: 100k agents, full year", flush=True)
    with S.Engine(max_slots=1) as e:
        e.set_day(256)
        slot = e.admit(mode=S.MODE_CONFIG, seed=w["seed"], horizon=steps, temp=temp, env=env,
                       max_agents=w["max_agents"], dmr=W.DMR, layout=layout, agents=agents)
        e.step_days(steps // 24)
        st, step, _ = e.status(slot)
        rows = e.rows(slot, 0, steps)
        lay, ag = e.export(slot)
        hist = e.histogram(slot)
    assert st == S.DONE and step == steps
    check_against(fx, rows, lay, ag, hist, "engine, day launches")


def test_day_refuses_more_work_groups_than_chunks_allow():
    with S.Engine(max_slots=1) as e:
        with pytest.raises(S.SampoError):
            e.set_day(257)
        e.set_day(0)
