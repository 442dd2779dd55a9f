"""GPU parity of the species plug-in: engines and the icl driver loop run a
species other than include/properties.h, against the oracle given the same
table.  Bar: bit-exact rows, ranges, agents and histogram."""
import numpy as np
import pytest

import sampo_amd as S
from oracle import pyoracle as O
from test_gpu_parity import agents_equal, intervention_env, rows_equal

pytestmark = pytest.mark.gpu


def hours(*ranges):
    return sum(1 << h for a, b in ranges for h in range(a, b + 1))


def variant_a():
    """Every member moved off its properties.h value."""
    sp = S.default_species().copy()
    sp["egg_incubation"] = (-0.8, 55.0)
    sp["hatch_threshold"] = (0.3, 0.6, 0.8, 0.95)
    sp["hatch_slope"] = (30.0, 50.0, 200.0, 400.0, 1000.0)
    sp["hatch_offset"] = (1.0, -5.0, -100.0, -250.0, -800.0)
    sp["larva_delay"] = (1.3, 0.2)
    sp["dev_temp_range"] = (10.0, 40.0)
    sp["dev_rates"] = (S.default_species()["dev_rates"] * np.float32(1.3)).astype(np.float32)
    sp["pupa_delay"] = (-0.7, 45.5)
    sp["immature_delay"] = (-2.0, 90.0)
    sp["mate_seeking_delay"] = 2.0
    sp["bms_delay"] = 1.0
    sp["bmd_delay"] = (-1.0, 60.0)
    sp["egg_batch"] = (120.0, 25.0, 0.85)
    sp["lay_probability"] = 0.4
    sp["bite_human"], sp["bite_indoor"], sp["rest_indoor"] = 0.7, 0.8, 0.6
    sp["window"] = (hours((0, 23)), hours((0, 8), (17, 23)), hours((0, 5), (19, 23)), hours((6, 20)),
                    hours((17, 19)), hours((0, 6), (18, 23)), hours((0, 23)), hours((0, 4), (20, 23)))
    assert S.check_species(sp)
    return sp


def variant_b():
    """A shorter, hotter cycle: small delays, a wide mate-seeking window."""
    sp = S.default_species().copy()
    sp["immature_delay"] = (-1.5, 70.0)
    sp["bmd_delay"] = (-0.9, 50.0)
    sp["egg_batch"] = (90.0, 40.0, 0.9)
    sp["window"][4] = hours((16, 22))
    sp["bite_human"] = 0.35
    assert S.check_species(sp)
    return sp


def test_engine_species_reference_mode_matches_oracle():
    days, seed = 24, 41
    steps = 24 * days
    temp = S.gen_temperature(seed, days, mean=26.0, amp=9.0, noise=3.0)
    env = intervention_env(steps, 9000.0, seed)
    sp = variant_a()
    sim = O.Sim(temp=temp, env=env, capacity=400_000, seed=seed, n_eggs=4000, species=sp)
    orows = sim.run(steps)
    assert orows["count"][-1][3:].sum() > 0 and orows["bites"].sum() > 0 and orows["numCycles"].sum() > 0
    with S.Engine(max_slots=1) as e:
        slot = e.admit(mode=S.MODE_REFERENCE, seed=seed, horizon=steps, temp=temp, env=env, n_eggs=4000,
                       max_agents=400_000, species=sp)
        e.step(steps)
        rows_equal(orows, e.rows(slot, 0, steps))
        lay_o, ag_o, _ = sim.export()
        lay_g, ag_g = e.export(slot)
        assert np.array_equal(lay_o, lay_g)
        agents_equal(ag_o, ag_g)
        assert np.array_equal(sim.histogram(), e.histogram(slot))


def test_engine_mixed_species_batch_matches_oracle():
    """One launch advances instances of different species: each equals its
    own oracle run."""
    days = 14
    steps = 24 * days
    cases = []
    for i, sp in enumerate((None, variant_a(), variant_b())):
        seed = 50 + i
        temp = S.gen_temperature(seed, days, mean=27.0, amp=4.0)
        lay, ag = S.gen_population(seed, 6000, float(temp[0]))
        env = intervention_env(steps, 30000.0, seed)
        cases.append((seed, temp, lay, ag, env, sp))
    dmr = (0.05, 0.03, 0.02, 0.1)
    with S.Engine(max_slots=4) as e:
        slots = [e.admit(mode=S.MODE_CONFIG, seed=seed, horizon=steps, temp=temp, env=env, max_agents=300_000,
                         dmr=dmr, layout=lay, agents=ag, species=sp)
                 for seed, temp, lay, ag, env, sp in cases]
        e.step(steps)
        for slot, (seed, temp, lay, ag, env, sp) in zip(slots, cases):
            sim = O.Sim(temp=temp, env=env, capacity=300_000, seed=seed, layout=lay, agents=ag, mode=1, dmr=dmr,
                        species=sp)
            rows_equal(sim.run(steps), e.rows(slot, 0, steps))
            lay_o, ag_o, _ = sim.export()
            lay_g, ag_g = e.export(slot)
            assert np.array_equal(lay_o, lay_g)
            agents_equal(ag_o, ag_g)


def test_engine_rejects_invalid_species():
    sp = S.default_species().copy()
    sp["hatch_threshold"] = (0.9, 0.5, 0.95, 0.99)
    with S.Engine(max_slots=1) as e:
        with pytest.raises(S.SampoError):
            e.admit(mode=S.MODE_REFERENCE, seed=1, horizon=24, temp=np.full(24, 27.0, np.float32),
                    env=S.make_env(), n_eggs=100, max_agents=10_000, species=sp)


def test_icl_driver_species_file_matches_oracle(tmp_path):
    """The speciesHeader argument of the driver (-DSPECIES in every kernel's
    build options) names a species file; initAgents and updateAgents read
    the table it holds."""
    steps, seed = 24 * 16, 43
    temp = S.gen_temperature(seed, 16, mean=27.0, amp=6.0)
    env = intervention_env(steps, 9000.0, seed)
    sp = variant_a()
    path = tmp_path / "variant_a.species"
    S.write_species(path, sp)
    sim = O.Sim(temp=temp, env=env, capacity=400_000, seed=seed, n_eggs=3000, species=sp)
    orows = sim.run(steps)
    g = S.run_abms(seed=seed, n_eggs=3000, max_agents=400_000, steps=steps, temp=temp, env=env,
                   species=str(path))
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
    assert np.array_equal(g["histogram"], sim.histogram())
    # properties.h by name is the default table
    g2 = S.run_abms(seed=seed, n_eggs=3000, max_agents=400_000, steps=24, temp=temp, env=env,
                    species="include/properties.h", export=False)
    sim2 = O.Sim(temp=temp[:24], env=env[:24], capacity=400_000, seed=seed, n_eggs=3000)
    rows_equal(sim2.run(24), g2["rows"])
