"""GPU parity: the HIP paths against the CPU oracle on the same inputs.

Bar: bit-exact.  Every printStats row, the final ranges, every field of
every agent in reference index order, and the age histogram must equal the
oracle's.  Sizes are chosen so the oracle finishes in seconds.
"""
import numpy as np
import pytest

import sampo_amd as S
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

ROWF = ("l1de", "females", "potInfective", "numCycles", "sumCycles", "bites", "infectBites")


def rows_equal(orows, grows):
    assert len(orows) == len(grows)
    assert np.array_equal(orows["count"], grows["count"])
    for f in ROWF:
        assert np.array_equal(orows[f], grows[f]), f


def agents_equal(a, b):
    """Field by field, as bits (a -0 is not a +0, a NaN equals itself)."""
    assert a.shape == b.shape
    for f in a.dtype.names:
        u = f"u{a[f].dtype.itemsize}"
        assert np.array_equal(np.ascontiguousarray(a[f]).view(u), np.ascontiguousarray(b[f]).view(u)), f


def ref_inputs(steps, cc, temp=30.0):
    return np.full(steps, temp, np.float32), S.make_env(cc=cc)


@pytest.mark.parametrize("n_eggs,cc,days,seed", [(3200, 16000.0, 12, 1), (500, 3000.0, 20, 7), (64, 800.0, 6, 3)])
def test_engine_reference_mode_matches_oracle(n_eggs, cc, days, seed):
    steps = 24 * days
    temp, env = ref_inputs(steps, cc)
    sim = O.Sim(temp=temp, env=env, capacity=400_000, seed=seed, n_eggs=n_eggs)
    orows = sim.run(steps)
    with S.Engine(max_slots=2) as e:
        slot = e.admit(mode=S.MODE_REFERENCE, seed=seed, horizon=steps, temp=temp, env=env, n_eggs=n_eggs,
                       max_agents=400_000)
        e.step(steps)
        st, step, n = e.status(slot)
        assert st == S.DONE and step == steps
        rows_equal(orows, e.rows(slot, 0, steps))
        lay_o, ag_o, _ = sim.export()
        lay_g, ag_g = e.export(slot)
        assert np.array_equal(lay_o, lay_g)
        agents_equal(ag_o, ag_g)
        assert np.array_equal(sim.histogram(), e.histogram(slot))


def config_inputs(n0, days, seed, mean=25.0, amp=5.0):
    temp = S.gen_temperature(seed, days, mean=mean, amp=amp)
    lay, ag = S.gen_population(seed, n0, float(temp[0]))
    return temp, lay, ag


@pytest.mark.parametrize("n0,days,seed,dmr", [(20000, 10, 11, (0.05, 0.03, 0.02, 0.1)),
                                               (3000, 15, 12, (0.04, 0.045, 0.01, 0.12))])
def test_engine_config_mode_matches_oracle(n0, days, seed, dmr):
    steps = 24 * days
    temp, lay, ag = config_inputs(n0, days, seed)
    env = S.make_env(cc=5.0 * n0)
    sim = O.Sim(temp=temp, env=env, capacity=60 * n0, seed=seed, layout=lay, agents=ag, mode=1, dmr=dmr)
    orows = sim.run(steps)
    with S.Engine(max_slots=2) as e:
        slot = e.admit(mode=S.MODE_CONFIG, seed=seed, horizon=steps, temp=temp, env=env, max_agents=60 * n0,
                       dmr=dmr, layout=lay, agents=ag)
        e.step(steps)
        rows_equal(orows, e.rows(slot, 0, steps))
        lay_o, ag_o, _ = sim.export()
        lay_g, ag_g = e.export(slot)
        assert np.array_equal(lay_o, lay_g)
        agents_equal(ag_o, ag_g)


@pytest.mark.parametrize("field", ["csd", "humanBloodmealCount", "aux", "cycleLength"])
def test_engine_full_record_when_immatures_carry_adult_words(field):
    """Initial eggs, larvae and pupae that carry words an immature keeps at
    zero (here: a development sum, blood meals, an egg's larval delay sum or
    a cycle length) make the instance store every word in every state; the
    run still equals the oracle, including those agents' own fields."""
    n0, days, seed, dmr = 4000, 6, 21, (0.04, 0.03, 0.02, 0.1)
    steps = 24 * days
    temp, lay, ag = config_inputs(n0, days, seed)
    rng = np.random.default_rng(3)
    n_imm = int(lay[2][1] - lay[2][0] + lay[1][1] - lay[1][0] + lay[0][1] - lay[0][0])
    pick = rng.choice(n_imm, 200, replace=False)
    ag = ag.copy()
    if field == "csd":
        ag["csd"][pick] = rng.uniform(0.0, 2.0, 200).astype(np.float32)
    elif field == "aux":  # only an egg's larval delay sum is zero by the rules
        pick = pick[pick < int(lay[0][1])]
        ag["aux"][pick] = np.float32(0.25).view(np.uint32)
    else:
        ag[field][pick] = rng.integers(1, 5, 200)
    env = S.make_env(cc=5.0 * n0)
    sim = O.Sim(temp=temp, env=env, capacity=60 * n0, seed=seed, layout=lay, agents=ag, mode=1, dmr=dmr)
    orows = sim.run(steps)
    with S.Engine(max_slots=1) as e:
        slot = e.admit(mode=S.MODE_CONFIG, seed=seed, horizon=steps, temp=temp, env=env, max_agents=60 * n0,
                       dmr=dmr, layout=lay, agents=ag)
        l0, a0 = e.export(slot)
        agents_equal(ag, a0)  # stored and read back unchanged before any iteration
        e.step(steps)
        rows_equal(orows, e.rows(slot, 0, steps))
        lay_o, ag_o, _ = sim.export()
        lay_g, ag_g = e.export(slot)
        assert np.array_equal(lay_o, lay_g)
        agents_equal(ag_o, ag_g)


def test_engine_development_sums_of_minus_zero_keep_their_sign():
    """Adults that start with a development sum of -0 keep it while it does
    not grow: the reference adds cond * increment in binary64, which is -0
    below 16 degC (the increment is negative there) and +0 above, so an
    adult's -0 stays -0 on cold hours and turns +0 on warm ones.  The run
    crosses 16 degC both ways; the exported sums are compared as bits."""
    n0, days, seed, dmr = 4000, 4, 23, (0.04, 0.03, 0.02, 0.1)
    steps = 24 * days
    temp = (15.5 - 3.0 * np.sin(np.arange(steps) * (2.0 * np.pi / 24.0))).astype(np.float32)
    assert (temp < 16.0).any() and (temp > 16.0).any()
    lay, ag = S.gen_population(seed, n0, float(temp[0]))
    n_imm = sum(int(lay[s][1] - lay[s][0]) for s in range(3))
    ag = ag.copy()
    ag["csd"][n_imm:] = np.float32(-0.0)
    env = S.make_env(cc=5.0 * n0)
    sim = O.Sim(temp=temp, env=env, capacity=60 * n0, seed=seed, layout=lay, agents=ag, mode=1, dmr=dmr)
    with S.Engine(max_slots=1) as e:
        slot = e.admit(mode=S.MODE_CONFIG, seed=seed, horizon=steps, temp=temp, env=env, max_agents=60 * n0,
                       dmr=dmr, layout=lay, agents=ag)
        signs, orows = [], []
        for k in range(steps):
            orows.append(sim.run(1))
            e.step(1)
            lay_o, ag_o, _ = sim.export()
            lay_g, ag_g = e.export(slot)
            assert np.array_equal(lay_o, lay_g)
            agents_equal(ag_o, ag_g)
            signs.append(int(np.count_nonzero(np.signbit(ag_g["csd"]))))
        rows_equal(np.concatenate(orows), e.rows(slot, 0, steps))
    assert max(signs) > 0 and min(signs) < max(signs)  # some -0 sums kept, some turned


def test_engine_refuses_non_finite_temperatures():
    temp, lay, ag = config_inputs(500, 2, 5)
    temp = temp.copy()
    temp[30] = np.nan
    with S.Engine(max_slots=1) as e:
        with pytest.raises(S.SampoError):
            e.admit(mode=S.MODE_CONFIG, seed=5, horizon=48, temp=temp, env=S.make_env(cc=2500.0), max_agents=20000,
                    layout=lay, agents=ag)


def intervention_env(steps, cc, seed):
    """A time-varying environment with every intervention switched on in
    some hours: larvicide, ITN, IRS and ovitrap coverage x effectiveness
    (updateAgents.cl:90-97, 166-173, 200-207, 236-264)."""
    rng = np.random.default_rng(seed)
    env = np.repeat(S.make_env(cc=cc), steps)
    env["bloodmealSuccess"] = rng.uniform(0.2, 0.9, steps).astype(np.float32)
    for f, hi in (("ITNValue", 0.4), ("IRSValue", 0.2), ("larvacideValue", 0.3), ("oviTrapValue", 0.5)):
        v = rng.uniform(0.0, hi, steps).astype(np.float32)
        v[rng.uniform(size=steps) < 0.3] = 0.0  # some hours without the intervention
        env[f] = v
    return env


@pytest.mark.parametrize("mode", ["reference", "config"])
def test_engine_interventions_match_oracle(mode):
    days, seed = (20 if mode == "reference" else 12), 23
    steps = 24 * days
    if mode == "reference":
        temp = np.full(steps, 29.0, np.float32)
        env = intervention_env(steps, 12000.0, seed)
        kw = dict(n_eggs=4000)
        akw = dict(mode=S.MODE_REFERENCE, n_eggs=4000)
    else:
        temp, lay, ag = config_inputs(8000, days, seed)
        env = intervention_env(steps, 0.025 * 8000, seed)
        dmr = (0.05, 0.03, 0.02, 0.1)
        kw = dict(layout=lay, agents=ag, mode=1, dmr=dmr)
        akw = dict(mode=S.MODE_CONFIG, dmr=dmr, layout=lay, agents=ag)
    sim = O.Sim(temp=temp, env=env, capacity=400_000, seed=seed, **kw)
    orows = sim.run(steps)
    assert orows["count"][-1].sum() > 0
    with S.Engine(max_slots=1) as e:
        slot = e.admit(seed=seed, horizon=steps, temp=temp, env=env, max_agents=400_000, **akw)
        e.step(steps)
        rows_equal(orows, e.rows(slot, 0, steps))
        lay_o, ag_o, _ = sim.export()
        lay_g, ag_g = e.export(slot)
        assert np.array_equal(lay_o, lay_g)
        agents_equal(ag_o, ag_g)


def test_icl_driver_interventions_match_oracle():
    steps, seed = 24 * 8, 29
    temp = np.full(steps, 28.0, np.float32)
    env = intervention_env(steps, 9000.0, seed)
    sim = O.Sim(temp=temp, env=env, capacity=400_000, seed=seed, n_eggs=3000)
    orows = sim.run(steps)
    g = S.run_abms(seed=seed, n_eggs=3000, max_agents=400_000, steps=steps, temp=temp, env=env, export=False)
    assert g["status"] == 0 and g["steps_done"] == steps
    rows_equal(orows, g["rows"])
    assert np.array_equal(sim.layout(), g["layout"])


def test_engine_hip_graph_day_equals_plain_steps():
    steps = 24 * 6
    temp, env = ref_inputs(steps, 8000.0)
    out = []
    for graph in (False, True):
        with S.Engine(max_slots=1) as e:
            slot = e.admit(seed=5, horizon=steps, temp=temp, env=env, n_eggs=1600, max_agents=1_000_000)
            if graph:
                e.step_days(steps // 24)
            else:
                e.step(steps)
            out.append((e.rows(slot, 0, steps + 1), e.export(slot)[1]))
    rows_equal(out[0][0], out[1][0])
    agents_equal(out[0][1], out[1][1])


def test_engine_captured_iterations_equal_plain_steps():
    """sampo_engine_capture/replay (one graph of n iterations with events
    around each pass) gives the plain iterations' results."""
    steps = 24 * 3 + 5
    temp, lay, ag = config_inputs(5000, 4, 31)
    env = S.make_env(cc=5.0 * 5000)
    dmr = (0.05, 0.03, 0.02, 0.1)
    out = []
    for captured in (False, True):
        with S.Engine(max_slots=1) as e:
            slot = e.admit(mode=S.MODE_CONFIG, seed=31, horizon=24 * 4, temp=temp, env=env, max_agents=200_000,
                           dmr=dmr, layout=lay, agents=ag)
            e.step(7)
            if captured:
                e.capture(steps)
                e.replay()
                wall, pss = e.replay_times()
                assert 0.0 < pss <= wall
            else:
                e.step(steps)
            e.sync()
            out.append((e.rows(slot, 0, 7 + steps), e.export(slot)))
    rows_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1][0], out[1][1][0])
    agents_equal(out[0][1][1], out[1][1][1])


def test_engine_batched_instances_equal_solo_runs():
    """Instances of different sizes, seeds and modes in one launch per
    iteration are each bit-identical to running alone."""
    steps = 24 * 5
    specs = []
    for i, (n, cc) in enumerate([(3000, 15000.0), (200, 1000.0), (9000, 40000.0), (40, 300.0)]):
        temp, env = ref_inputs(steps, cc, temp=26.0 + i)
        specs.append(dict(mode=S.MODE_REFERENCE, seed=100 + i, horizon=steps, temp=temp, env=env, n_eggs=n,
                          max_agents=50 * n + 10000))
    temp, lay, ag = config_inputs(5000, 5, 21)
    specs.append(dict(mode=S.MODE_CONFIG, seed=21, horizon=steps, temp=temp, env=S.make_env(cc=25000.0),
                      max_agents=300000, dmr=(0.05, 0.03, 0.02, 0.1), layout=lay, agents=ag))
    solo = []
    for sp in specs:
        with S.Engine(max_slots=1) as e:
            s = e.admit(**sp)
            e.step(steps)
            solo.append((e.rows(s, 0, steps + 1), e.export(s)[1]))
    with S.Engine(max_slots=8) as e:
        slots = [e.admit(**sp) for sp in specs]
        e.step_days(steps // 24)
        for s, (r, a) in zip(slots, solo):
            rows_equal(r, e.rows(s, 0, steps + 1))
            agents_equal(a, e.export(s)[1])


def test_engine_many_instances_match_solo_runs():
    """150 instances in one launch (more than one wave of 64 in the pass's
    instance search and in the plan's scan), with horizons of 1 to 4 days so
    that finished instances hold no chunks between running ones: each is the
    same whether admitted first or last, and a sample equals its run alone."""
    n, days = 150, 4
    specs = []
    for i in range(n):
        d = 1 + i % days
        temp, lay, ag = config_inputs(200 + 37 * (i % 80), d, 300 + i)
        specs.append(dict(mode=S.MODE_CONFIG, seed=300 + i, horizon=24 * d, temp=temp,
                          env=S.make_env(cc=5.0 * (200 + 37 * (i % 80))), max_agents=40_000,
                          dmr=(0.05, 0.03, 0.02, 0.1), layout=lay, agents=ag))
    res = []
    for order in (range(n), range(n - 1, -1, -1)):
        with S.Engine(max_slots=n) as e:
            e.set_chunk(256, 1 << 20)
            slots = {i: e.admit(**specs[i]) for i in order}
            e.step_days(days)
            got = {}
            for i, s in slots.items():
                st, step, _ = e.status(s)
                assert st == S.DONE and step == specs[i]["horizon"], (i, st, step)
                got[i] = (e.rows(s, 0, specs[i]["horizon"] + 1), e.export(s)[1])
            res.append(got)
    for i in range(n):
        rows_equal(res[0][i][0], res[1][i][0])
        agents_equal(res[0][i][1], res[1][i][1])
    for i in range(0, n, 15):
        with S.Engine(max_slots=1) as e:
            e.set_chunk(256, 1 << 20)
            s = e.admit(**specs[i])
            e.step(specs[i]["horizon"])
            rows_equal(e.rows(s, 0, specs[i]["horizon"] + 1), res[0][i][0])
            agents_equal(e.export(s)[1], res[0][i][1])


@pytest.mark.parametrize("target", [8, 16, 64, 200])
def test_engine_chunks_of_three_sizes_match_oracle(target):
    """A small target chunk count over a minimum of 64 agents: half the
    chunks big with 80 % of the agents, an eighth of one size, the rest half
    that size (or equal chunks where rounding leaves no tail), as the pass
    cuts one large run."""
    n0, days, seed, dmr = 20000, 4, 33, (0.05, 0.03, 0.02, 0.1)
    steps = 24 * days
    temp, lay, ag = config_inputs(n0, days, seed)
    env = S.make_env(cc=0.05 * n0)
    sim = O.Sim(temp=temp, env=env, capacity=60 * n0, seed=seed, layout=lay, agents=ag, mode=1, dmr=dmr)
    orows = sim.run(steps)
    with S.Engine(max_slots=1) as e:
        e.set_chunk(64, target)
        slot = e.admit(mode=S.MODE_CONFIG, seed=seed, horizon=steps, temp=temp, env=env, max_agents=60 * n0,
                       dmr=dmr, layout=lay, agents=ag)
        e.step(steps)
        rows_equal(orows, e.rows(slot, 0, steps))
        lay_o, ag_o, _ = sim.export()
        lay_g, ag_g = e.export(slot)
        assert np.array_equal(lay_o, lay_g)
        agents_equal(ag_o, ag_g)


@pytest.mark.parametrize("chunk", [64, 200, 1000])
def test_engine_small_chunks_match_oracle(chunk):
    """Many chunks per state: blocks split across chunks, tiles that span
    states, runs moved in from the previous state in every chunk."""
    n0, days, seed, dmr = 6000, 6, 31, (0.05, 0.03, 0.02, 0.1)
    steps = 24 * days
    temp, lay, ag = config_inputs(n0, days, seed)
    env = S.make_env(cc=0.05 * n0)
    sim = O.Sim(temp=temp, env=env, capacity=60 * n0, seed=seed, layout=lay, agents=ag, mode=1, dmr=dmr)
    orows = sim.run(steps)
    with S.Engine(max_slots=1) as e:
        e.set_chunk(chunk, 1 << 20)
        slot = e.admit(mode=S.MODE_CONFIG, seed=seed, horizon=steps, temp=temp, env=env, max_agents=60 * n0,
                       dmr=dmr, layout=lay, agents=ag)
        e.step(steps)
        rows_equal(orows, e.rows(slot, 0, steps))
        lay_o, ag_o, _ = sim.export()
        lay_g, ag_g = e.export(slot)
        assert np.array_equal(lay_o, lay_g)
        agents_equal(ag_o, ag_g)
        assert np.array_equal(sim.histogram(), e.histogram(slot))


def test_engine_growth_is_transparent():
    """Instances admitted with only as many slots as their first population
    pause when the output would not fit, grow between days and resume; each
    result is still the oracle's, bit for bit."""
    days = 30  # both populations fall for ~11 days, then outgrow their first buffers
    steps = 24 * days
    specs, sims = [], []
    for i, n0 in enumerate((4000, 700)):
        temp, lay, ag = config_inputs(n0, days, 40 + i)
        env = S.make_env(cc=0.025 * n0)
        dmr = (0.05, 0.03, 0.02, 0.1)
        sim = O.Sim(temp=temp, env=env, capacity=40 * n0, seed=40 + i, layout=lay, agents=ag, mode=1, dmr=dmr)
        sims.append((sim.run(steps), sim.export()))
        specs.append(dict(mode=S.MODE_CONFIG, seed=40 + i, horizon=steps, temp=temp, env=env, max_agents=40 * n0,
                          phys_slots=n0, dmr=dmr, layout=lay, agents=ag))
    with S.Engine(max_slots=4) as e:
        slots = [e.admit(**sp) for sp in specs]
        grown, paused_seen, rounds = 0, False, 0
        while True:
            st, step, _ = e.poll(4)
            if all(st[s] == S.DONE for s in slots):
                break
            assert all(st[s] in (S.RUNNING, S.GROW) for s in slots), st
            paused_seen |= any(st[s] == S.GROW for s in slots)
            grown += e.service(1.0)  # grow only when an instance must
            e.step_days(1)
            rounds += 1
            assert rounds < 4 * days
        assert grown > 0 and paused_seen
        for s, (orows, (lay_o, ag_o, _)) in zip(slots, sims):
            rows_equal(orows, e.rows(s, 0, steps))
            lay_g, ag_g = e.export(s)
            assert np.array_equal(lay_o, lay_g)
            agents_equal(ag_o, ag_g)


def test_engine_tile_shapes_agree():
    steps = 24 * 4
    temp, env = ref_inputs(steps, 20000.0)
    ref = None
    for wg, apt, waves in ((256, 1, 0), (128, 2, 0), (64, 4, 0), (512, 2, 0), (256, 4, 0), (512, 1, 6), (768, 1, 6),
                           (1024, 1, 0), (256, 1, 8)):
        with S.Engine(max_slots=1) as e:
            e.set_tile(wg, apt, waves)
            s = e.admit(seed=9, horizon=steps, temp=temp, env=env, n_eggs=5000, max_agents=1_000_000)
            e.step(steps)
            got = (e.rows(s, 0, steps + 1), e.export(s)[1])
        if ref is None:
            ref = got
        else:
            rows_equal(ref[0], got[0])
            agents_equal(ref[1], got[1])


def test_engine_status_capacity_and_extinction():
    steps = 24 * 30
    temp, env = ref_inputs(steps, 20000.0)
    with S.Engine(max_slots=2) as e:
        s = e.admit(seed=1, horizon=steps, temp=temp, env=env, n_eggs=2000, max_agents=2200)
        e.step(steps)
        st, step, _ = e.status(s)
        assert st == S.CAPACITY and step < steps
        t2, lay, ag = config_inputs(400, 2, 5)
        s2 = e.admit(mode=S.MODE_CONFIG, seed=5, horizon=48, temp=t2, env=S.make_env(cc=2000.0),
                     max_agents=10000, dmr=(1.0, 1.0, 1.0, 1.0), layout=lay, agents=ag)
        e.step(48)
        st2, step2, _ = e.status(s2)
        assert st2 == S.EXTINCT and step2 < 48
    sim = O.Sim(temp=temp, env=env, capacity=2200, seed=1, n_eggs=2000)
    sim.run(steps)
    assert sim.status == 1 and sim.step == step


def icl_vs_oracle(n_eggs, cc, steps, seed, serial, config=None):
    cap = 400_000
    if config is None:
        temp, env = ref_inputs(steps, cc)
        kw = dict(n_eggs=n_eggs)
    else:
        n0, dmr = config
        temp, lay, ag = config_inputs(n0, (steps + 23) // 24, seed)
        env = S.make_env(cc=cc)
        kw = dict(mode=S.MODE_CONFIG, dmr=dmr, layout=lay, agents=ag)
    sim = O.Sim(temp=temp, env=env, capacity=cap, seed=seed, **kw)
    orows = sim.run(steps)
    g = S.run_abms(seed=seed, max_agents=cap, steps=steps, temp=temp, env=env, serial_scan=serial, **kw)
    assert g["status"] == 0 and g["steps_done"] == steps
    rows_equal(orows, g["rows"])
    lay_o = sim.layout()
    assert np.array_equal(lay_o, g["layout"])
    _, ag_o, st_o = sim.export()
    idx = np.concatenate([np.arange(lay_o[r][0], lay_o[r][1]) for r in range(8)])
    for f in ("availableEggs", "cycleLength", "humanBloodmealCount", "delay", "numEggBatches", "aux"):
        assert np.array_equal(g["agents"][f][idx], ag_o[f]), f
    for f in ("ageInHours", "hoursInState", "isFemale", "csd"):
        assert np.array_equal(g["ages"][f][idx], ag_o[f]), f
    assert np.array_equal(g["states"]["state"][idx], st_o)
    assert np.array_equal(g["histogram"], sim.histogram())


def test_icl_driver_configuration_mode_matches_oracle():
    """The plain driver loop with an explicit initial population and
    constant per-stage mortality (the killAgentsStaged kernel)."""
    icl_vs_oracle(0, 150.0, 24 * 8, 17, serial=False, config=(6000, (0.05, 0.03, 0.02, 0.1)))


def test_icl_driver_parallel_scan_matches_oracle():
    icl_vs_oracle(3200, 16000.0, 24 * 10, 1, serial=False)


def test_icl_driver_serial_scan_matches_oracle():
    icl_vs_oracle(300, 2000.0, 24 * 6, 4, serial=True)
