"""The species plug-in as data (CPU).

The reference picks its species at kernel build time: -DSPECIES=<header>
(src/abms.c:565-568), and ships one header, include/properties.h.  Here a
species is a table (sampo_species) read per instance.  These tests pin

  - the default table of the library and of the oracle against the numbers
    read from include/properties.h and kernel/updateAgents.cl
    (tests/golden/constants.json);
  - the species file reader;
  - the oracle's use of each table member, through closed forms and
    invariants that a wrong member, index or random stream would break.
"""
import json
import os
import re

import numpy as np
import pytest

import sampo_amd as S
from oracle import pyoracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CONST = json.load(open(os.path.join(GOLD, "constants.json")))


def linear(expr):
    """(a, b) of a reference expression 'a*temperature+b' (f suffixes optional)."""
    m = re.fullmatch(r"\(?\s*(-?[0-9.]+)f?\s*\*\s*temperature\s*\+\s*([0-9.]+)f?\s*\)?", expr.strip())
    assert m, expr
    return float(m.group(1)), float(m.group(2))


def expected_default():
    c, sp = CONST, CONST["species"]
    e = np.zeros(1, S.SPECIES_DTYPE)[0]
    e["egg_incubation"] = linear(c["eggsIncubation"]["expr"])
    e["hatch_threshold"] = c["eggsHatching"]["thresholds"]
    e["hatch_slope"] = c["eggsHatching"]["m"]
    e["hatch_offset"] = c["eggsHatching"]["b"]
    e["larva_delay"] = (c["larvaDelay"]["mean"], c["larvaDelay"]["sd"])
    lo, hi = sp["larvaDevelopment"]["clamp"]
    assert sp["larvaDevelopment"]["base"] == int(lo)  # the table starts at the clamp's lower end
    e["dev_temp_range"] = (lo, hi)
    e["dev_rates"] = c["devRates"]["values"]
    e["pupa_delay"] = linear(c["pupaDelay"]["expr"])  # unsuffixed: binary64 constants
    e["immature_delay"] = linear(c["immatureDelay"]["expr"])
    e["mate_seeking_delay"] = sp["mateSeekingDelay"]["value"]
    e["bms_delay"] = sp["bloodMealSeekingDelay"]["value"]
    e["bmd_delay"] = linear(c["bloodMealDigestingDelay"]["expr"])
    e["egg_batch"] = (c["generateEggs"]["mean"], c["generateEggs"]["sd"], sp["eggBatchDecay"]["value"])
    e["lay_probability"] = sp["layProbability"]["value"]
    e["bite_human"] = sp["BITE_HUMAN_PROBABILITY"]["value"]
    e["bite_indoor"] = sp["BITE_INDOOR_PROBABILITY"]["value"]
    e["rest_indoor"] = sp["REST_INDOOR_PROBABILITY"]["value"]
    e["window"] = [w["mask"] for w in sp["window"]]
    return e


def test_default_species_is_properties_h():
    want = expected_default()
    assert [w["fn"] for w in CONST["species"]["window"]][-1] == "gravidsEggLayTime"
    for name, got in (("library", S.default_species()), ("oracle", O.default_species())):
        g = np.frombuffer(got.tobytes(), S.SPECIES_DTYPE)[0]
        for f in S.SPECIES_DTYPE.names:
            assert np.array_equal(g[f], want[f]), (name, f, g[f], want[f])
    # the pupa coefficients are the binary64 values of the unsuffixed literals
    assert S.default_species()["pupa_delay"][0] == -0.923 and S.default_species()["pupa_delay"][1] == 60.923
    assert S.check_species(S.default_species())


def test_night_window_is_the_reference_predicate():
    night = [h for h in range(24) if (h <= 6 or h >= 18)]
    assert CONST["species"]["window"][1]["mask"] == sum(1 << h for h in night)
    assert CONST["species"]["window"][4]["mask"] == 1 << 18


def test_species_file_round_trip_and_overrides(tmp_path):
    d = S.default_species()
    p = tmp_path / "full.species"
    S.write_species(p, d)
    assert S.read_species(p).tobytes() == d.tobytes()
    q = tmp_path / "part.species"
    q.write_text("# a variant\nimmature_delay -2.0f, 100.0f  # comment\nwindow_mate_seeking 17-19\n"
                 "pupa_delay -0.5 40.25\nbite_human 0.5\n")
    v = S.read_species(q)
    assert v["immature_delay"][0] == np.float32(-2.0) and v["immature_delay"][1] == np.float32(100.0)
    assert v["window"][4] == (1 << 17) | (1 << 18) | (1 << 19)
    assert v["pupa_delay"][0] == -0.5 and v["bite_human"] == np.float32(0.5)
    for f in S.SPECIES_DTYPE.names:  # everything else stays properties.h
        if f not in ("immature_delay", "window", "pupa_delay", "bite_human"):
            assert np.array_equal(v[f], d[f]), f


@pytest.mark.parametrize("text", ["no_such_member 1\n", "larva_delay 1 2 3\n", "larva_delay 1\n",
                                  "window_larva 5-30\n", "hatch_threshold 0.9 0.5 0.95 0.99\n",
                                  "bite_human 1.5\n", "dev_temp_range 12.5 40\n", "egg_batch 170 -1 0.8\n"])
def test_species_file_errors(tmp_path, text):
    p = tmp_path / "bad.species"
    p.write_text(text)
    with pytest.raises(S.SampoError):
        S.read_species(p)


# ------------------------------------------------- the oracle's use of each member


def oracle_run(species, *, steps, n_eggs=1500, temp=27.0, cc=4000.0, env=None, seed=5):
    t = np.full(max(steps, 1), temp, np.float32)  # temp[0] also sets the initial eggs' delays
    env = env if env is not None else S.make_env(cc=cc)
    sim = O.Sim(temp=t, env=env, capacity=300000, seed=seed, n_eggs=n_eggs, species=species)
    rows = sim.run(steps)
    return sim, rows


def variant(**kw):
    sp = S.default_species().copy()
    for k, v in kw.items():
        sp[k] = v
    assert S.check_species(sp)
    return sp


def test_explicit_default_table_equals_no_table():
    a, ra = oracle_run(None, steps=24 * 20)
    b, rb = oracle_run(S.default_species(), steps=24 * 20)
    assert np.array_equal(ra, rb) and np.array_equal(a.export()[1], b.export()[1])


def test_hatching_bands_and_incubation_give_the_initial_egg_delays():
    # flat bands: eggsHatching == (uint)(offset + 0.5); delay == a*T + b + that
    sp = variant(hatch_slope=[0, 0, 0, 0, 0], hatch_offset=[3.2, 3.2, 3.2, 3.2, 3.2],
                 egg_incubation=[-1.0, 50.0])
    sim, _ = oracle_run(sp, steps=0, temp=20.0)
    _, ag, _ = sim.export()
    assert len(ag) == 1500 and np.all(ag["delay"] == np.float32(30.0 + 3.0))
    # per-band offsets: each egg's delay tells which band its U2 fell in
    sp = variant(hatch_slope=[0, 0, 0, 0, 0], hatch_offset=[0, 100, 200, 300, 400],
                 hatch_threshold=[0.2, 0.4, 0.6, 0.8], egg_incubation=[0.0, 0.0])
    sim, _ = oracle_run(sp, steps=0)
    _, ag, _ = sim.export()
    band = np.round(ag["delay"] / 100).astype(int)
    assert set(band) == {0, 1, 2, 3, 4}
    frac = np.bincount(band, minlength=5) / len(band)
    assert np.all(np.abs(frac - 0.2) < 0.05)  # U2 uniform on [0, 1)


def test_larva_delay_with_zero_sd_is_the_mean():
    sim, rows = oracle_run(variant(larva_delay=[7.5, 0.0]), steps=24 * 6)
    lay, ag, st = sim.export()
    larvae = ag[st == 2]
    assert len(larvae) > 100 and np.all(larvae["delay"] == np.float32(7.5))


def test_dev_rates_accumulate_per_hour():
    # constant development rate c: cumulativeLarvalDelay after k larval hours is k float additions of c
    c = np.float32(0.03125)  # exact in binary32, so k*c is exact too
    sim, rows = oracle_run(variant(dev_rates=[c] * 32, larva_delay=[1e9, 0.0]), steps=24 * 6)
    lay, ag, st = sim.export()
    larvae = ag[st == 2]
    assert len(larvae) > 100
    cum = larvae["aux"].view(np.float32)
    assert np.array_equal(cum, (larvae["hoursInState"] * c).astype(np.float32))
    # the index follows the clamp range: moving the table's start by one degree shifts the entry used
    rates = np.arange(1, 33, dtype=np.float32) / 1024
    sp = variant(dev_rates=rates, dev_temp_range=[20.0, 48.0], larva_delay=[1e9, 0.0])
    sim, _ = oracle_run(sp, steps=24 * 6, temp=27.5)
    _, ag, st = sim.export()
    larvae = ag[st == 2]
    assert np.array_equal(larvae["aux"].view(np.float32), (larvae["hoursInState"] * rates[7]).astype(np.float32))


def test_closed_mate_seeking_window_stops_the_cycle():
    sim, rows = oracle_run(variant(window=[0xffffff, 0xfc007f, 0xfc007f, 0xffffff, 0, 0xfc007f, 0xffffff,
                                           0xfc007f]), steps=24 * 30)
    assert rows["count"][:, 4].max() > 0  # mate seekers appear ...
    assert rows["count"][:, 5:].sum() == 0 and rows["bites"].sum() == 0  # ... and never leave


def test_delays_after_transitions_follow_the_table():
    sp = variant(immature_delay=[0.0, 30.0], mate_seeking_delay=3.0, bms_delay=2.0, bmd_delay=[0.0, 9.0],
                 pupa_delay=[0.0, 11.0])
    want = {4: 11.0, 8: 30.0, 16: 3.0, 32: 2.0, 64: 9.0}  # pupa, immature, mate seeking, BMS, BMD
    seen = set()
    sim = O.Sim(temp=np.full(24 * 40, 27.0, np.float32), env=S.make_env(cc=4000.0), capacity=300000, seed=5,
                n_eggs=1500, species=sp)
    for day in range(40):  # every agent in these states got its delay from the table on entry
        sim.run(24)
        _, ag, st = sim.export()
        for s_, d in want.items():
            x = ag[st == s_]
            if len(x):
                seen.add(s_)
                assert np.all(x["delay"] == np.float32(d)), (day, s_)
    assert seen == set(want)


def test_egg_batch_without_noise_or_decay():
    sim, rows = oracle_run(variant(egg_batch=[150.0, 0.0, 1.0]), steps=24 * 25)
    _, ag, st = sim.export()
    fresh = ag[(st == 128) & (ag["aux"] == 0)]  # gravids that have not tried to lay yet
    assert len(fresh) > 0 and np.all(fresh["availableEggs"] == 150)
    assert rows["count"][:, 7].max() > 0


def test_lay_probability_zero_and_one():
    _, r0 = oracle_run(variant(lay_probability=0.0), steps=24 * 25)
    assert r0["numCycles"].sum() == 0  # no gravid ever lays
    _, r1 = oracle_run(variant(lay_probability=1.0), steps=24 * 25)
    _, rd = oracle_run(None, steps=24 * 25)
    assert r1["numCycles"].sum() > rd["numCycles"].sum() > 0


def one_cycle_run(sp, env, days=30, n0=6000, seed=11):
    """Configuration mode without mortality and with laying switched off:
    every female that reaches BMS bites once, stays in BMD or gravid, and
    keeps her humanBloodmealCount (0 or 1)."""
    steps = 24 * days
    temp = np.full(steps, 27.0, np.float32)
    lay, ag = S.gen_population(seed, n0, 27.0)
    sim = O.Sim(temp=temp, env=env, capacity=4 * n0 + 4096, seed=seed, layout=lay, agents=ag, mode=1,
                dmr=(0, 0, 0, 0), species=sp)
    rows = sim.run(steps)
    _, a, st = sim.export()
    return rows, a[(st == 64) | (st == 128)]


@pytest.mark.parametrize("bh", [0.5, 0.8])
def test_bite_human_probability_uses_its_own_stream(bh):
    """A fed female bit a human with probability BITE_HUMAN_PROBABILITY when
    the draw is a stream of its own (seeds[3], updateAgents.cl:175).  With
    ITN present the survivors of the ITN draw have U2 > ITN x BH, so reusing
    U2 would give (BH - ITN x BH) / (1 - ITN x BH) instead, and reusing U1
    (the blood-meal draw, U1 <= 0.25) would give 1."""
    itn = 0.6
    rows, fed = one_cycle_run(variant(bite_human=bh, lay_probability=0.0), S.make_env(cc=1e9, itn=itn))
    assert len(fed) > 1500 and set(np.unique(fed["humanBloodmealCount"])) <= {0, 1}
    assert rows["bites"].sum() == len(fed)  # every bite is a fed, living female
    p = fed["humanBloodmealCount"].mean()
    sd = np.sqrt(bh * (1 - bh) / len(fed))
    assert abs(p - bh) < 4 * sd, (p, bh)
    k = itn * bh
    assert abs(p - (bh - k) / (1 - k)) > 8 * sd and abs(p - 1.0) > 8 * sd
    # and the ITN kill probability is ITN x BH x BITE_INDOOR: fraction killed among those that tried
    attempts = len(fed) + (rows["count"][0].sum() - rows["count"][-1].sum())  # mortality is 0: losses are ITN
    killed = attempts - len(fed)
    q = killed / attempts
    assert abs(q - k) < 4 * np.sqrt(k * (1 - k) / attempts), (q, k)


def test_bite_human_zero_means_no_infection():
    sim, rows = oracle_run(variant(bite_human=0.0), steps=24 * 40, temp=30.0)
    _, ag, _ = sim.export()
    assert rows["bites"].sum() > 0
    assert ag["humanBloodmealCount"].sum() == 0 and np.all(ag["csd"] == 0)
    assert rows["infectBites"].sum() == 0 and rows["potInfective"].sum() == 0


def test_rest_and_bite_indoor_scale_the_interventions():
    env = S.make_env(cc=4000.0, irs=0.4, itn=0.4)
    _, r_off = oracle_run(variant(rest_indoor=0.0, bite_indoor=0.0), steps=24 * 30, env=env)
    _, r_none = oracle_run(None, steps=24 * 30, env=S.make_env(cc=4000.0))
    assert np.array_equal(r_off, r_none)  # the draws are skipped when the products are 0
