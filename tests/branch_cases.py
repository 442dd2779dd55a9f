"""Closed-form checks of every update branch on hand-built populations
(test helper, not a test module).

Each case is a handful of agents placed in chosen states of the reference
layout (each range starts at the next multiple of 64 after the previous
end), run with zero per-stage mortality in configuration mode, and a check
that recomputes what the reference's kernels (kernel/updateAgents.cl,
include/properties.h, kernel/initAgents.cl) must produce for them from
  - the random draws of tests/kat_rng.py (numpy HybridTaus from the constants
    parsed out of include/gpuRand.h, seed sets from the host C library), at
    each agent's padded index, and
  - the model constants parsed into tests/golden/constants.json,
evaluated here in numpy binary32/binary64 in the reference's operation order.
None of it calls the oracle's or the product's code.  A case's random seed is
chosen, from the same draws, so that both outcomes of each random branch
occur among its agents.

run(case, runner) returns (rows, snaps): rows[t] is the printStats row at
the start of iteration t (t = 0..steps), snaps[t] the (layout, agents) at the
start of iteration t.  The CPU tests run the cases through the oracle, the
GPU tests through the engine; both use the same checks.
"""
import json
import os
from dataclasses import dataclass, field

import numpy as np

from tests import kat_rng as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONST = json.load(open(os.path.join(ROOT, "tests", "golden", "constants.json")))

AGENT = np.dtype([("ageInHours", "<f4"), ("hoursInState", "<f4"), ("delay", "<f4"), ("csd", "<f4"),
                  ("aux", "<u4"), ("availableEggs", "<u4"), ("cycleLength", "<u4"),
                  ("humanBloodmealCount", "<u4"), ("numEggBatches", "<u4"), ("isFemale", "<u4")])
ENV = np.dtype([("bloodmealSuccess", "<f4"), ("ITNValue", "<f4"), ("IRSValue", "<f4"),
                ("carryingCapacity", "<f4"), ("larvacideValue", "<f4"), ("oviTrapValue", "<f4")])
EGG, LARVA, PUPA, IMM, MS, BMS, BMD, GRAVID = range(8)
F32 = np.float32


def env(bloodmeal=0.25, itn=0.0, irs=0.0, cc=160000.0, larvicide=0.0, ovitrap=0.0):
    return np.array([(bloodmeal, itn, irs, cc, larvicide, ovitrap)], dtype=ENV)


def population(groups):
    """(layout (8, 2), dense agents) of {state: [field dicts]} in the
    reference layout."""
    lay = np.zeros((8, 2), np.uint32)
    n = sum(len(v) for v in groups.values())
    agents = np.zeros(n, AGENT)
    end, k = 0, 0
    for s in range(8):
        start = 0 if s == 0 else ((end + 64) // 64) * 64
        items = groups.get(s, [])
        lay[s] = (start, start + len(items))
        end = start + len(items)
        for it in items:
            for f, v in it.items():
                agents[f][k] = v
            k += 1
    return lay, agents


def states_of(layout):
    return np.concatenate([np.full(int(b - a), s, np.int64) for s, (a, b) in enumerate(layout)])


def gids_of(layout):
    return np.concatenate([np.arange(int(a), int(b), dtype=np.int64) for a, b in layout])


# ------------------------------------------------------------ model pieces

def dev_rate(t):
    lo, hi = CONST["species"]["larvaDevelopment"]["clamp"]
    base = CONST["species"]["larvaDevelopment"]["base"]
    c = min(max(F32(t), F32(lo)), F32(hi))
    return F32(CONST["devRates"]["values"][min(int(c) - base, 31)])


def egg_incubation(t):   # (-0.923f*T + 60.923f), binary32
    return F32(F32(-0.923) * F32(t)) + F32(60.923)


def pupa_delay(t):       # -0.923*T + 60.923 in binary64 (unsuffixed literals), stored as float
    return F32(-0.923 * np.float64(F32(t)) + 60.923)


def immature_delay(t):   # -2.667f * T + 120.0f
    return F32(F32(-2.667) * F32(t)) + F32(120.0)


def bmd_delay(t):        # -1.231f * T + 77.231f
    return F32(F32(-1.231) * F32(t)) + F32(77.231)


def eggs_hatching(p):
    h = CONST["eggsHatching"]
    i = next((j for j, thr in enumerate(h["thresholds"]) if F32(p) < F32(thr)), len(h["thresholds"]))
    v = F32(F32(F32(h["m"][i]) * F32(p)) + F32(h["b"][i])) + F32(0.5)
    return F32(np.floor(v)) if v >= 1 else F32(0.0)


def box_muller64(u0, u1, mean, sd):
    """BoxMuller in binary64 from the binary32 angle the reference forms
    (theta = 2 * M_PI_F * u1 in binary32)."""
    theta = np.float64(F32(F32(2.0) * F32(np.pi)) * F32(u1))
    return np.sqrt(-2.0 * np.log(np.float64(u0))) * np.sin(theta) * sd + mean


def csd_increment(t):
    """updateAgents.cl:69-70: 1.0f / ((111.0f / (T - 16.0f)) * 14.0), the
    quotient in binary32 and the rest in binary64."""
    with np.errstate(divide="ignore"):
        q = F32(111.0) / (F32(t) - F32(16.0))
        return 1.0 / (np.float64(q) * 14.0)


# ------------------------------------------------------------ the cases

@dataclass
class Case:
    name: str
    groups: dict
    steps: int
    seed: int
    temp: np.ndarray
    env: np.ndarray
    check: object
    species: dict = field(default_factory=dict)  # members changed from include/properties.h


def on_a_draw(u):
    """A threshold equal to one of the draws (the median), so that the
    comparison's strictness decides that agent's outcome."""
    return F32(np.sort(np.asarray(u, dtype=np.float32))[len(u) // 2])


def draws(seed, t, k, gids):
    return K.hybrid(K.seed_sets(seed, t + 1)[t][k], np.asarray(gids, dtype=np.uint64))


def pick_seed(gids, k, pred, t=0, start=1):
    """The first seed whose draws of stream k at iteration t give both
    outcomes of pred over gids (and no exact zero: zero mortality still
    kills on a zero draw)."""
    gids = np.asarray(gids)
    for seed in range(start, start + 500):
        u = draws(seed, t, k, gids)
        u0 = draws(seed, t, 0, gids)
        p = pred(u)
        if p.any() and not p.all() and (u0 > 0).all():
            return seed
    raise AssertionError("no seed mixes the outcomes")


def steady_temp(steps, t=27.3):
    return np.full(steps + 1, t, np.float32)


def case_egg_hatch():
    groups = {EGG: [dict(ageInHours=5, hoursInState=5, delay=0.0, isFemale=1),
                    dict(ageInHours=3, hoursInState=3, delay=100.0)]}

    def check(rows, snaps):
        seed = c.seed
        lay1, a1 = snaps[1]
        st1 = states_of(lay1)
        # the ready egg (padded index 0) hatched; the other (index 1) aged
        lv = a1[st1 == LARVA]
        eg = a1[st1 == EGG]
        assert len(lv) == 1 and len(eg) == 1
        u1, u2 = draws(seed, 0, 1, [0])[0], draws(seed, 0, 2, [0])[0]
        want = box_muller64(u1, u2, CONST["larvaDelay"]["mean"], CONST["larvaDelay"]["sd"])
        assert lv["ageInHours"][0] == 24.0 and lv["hoursInState"][0] == 0.0
        assert abs(float(lv["delay"][0]) - want) <= 4e-7 * max(1.0, abs(want)), (lv["delay"][0], want)
        assert lv["aux"][0] == 0 and lv["isFemale"][0] == 1
        assert eg["ageInHours"][0] == 4.0 and eg["hoursInState"][0] == 4.0
        # one hour of larval development later
        lay2, a2 = snaps[2]
        lv2 = a2[states_of(lay2) == LARVA]
        assert lv2["aux"].view(np.float32)[0] == F32(0.0) + dev_rate(27.3)
        assert lv2["ageInHours"][0] == 25.0 and lv2["hoursInState"][0] == 1.0

    c = Case("egg_hatch", groups, 2, 1, steady_temp(2), env(), check)
    return c


def case_larva_pupate():
    d = dev_rate(27.3)
    acc = [F32(0.0)]
    for _ in range(25):
        acc.append(acc[-1] + d)
    d1, d2 = acc[3], acc[9]  # crossed on the 3rd add (hour 2, open) and the 9th (hour 8, closed until 18)
    groups = {LARVA: [dict(ageInHours=30, hoursInState=6, delay=d1),
                      dict(ageInHours=40, hoursInState=2, delay=d2)]}

    def check(rows, snaps):
        for t in range(1, c.steps + 1):
            lay, a = snaps[t]
            st = states_of(lay)
            # agent 1: larva through iteration 1, pupa from iteration 2 on
            if t <= 2:
                assert st[0] == LARVA and a["aux"].view(np.float32)[0] == acc[t]
            elif t <= 19:
                pu = a[st == PUPA][-1:]
                assert pu["hoursInState"][0] == t - 3
                assert pu["delay"][0] == pupa_delay(27.3) and pu["ageInHours"][0] == 30 + t
                assert pu["aux"].view(np.float32)[0] == acc[3]
            lv = a[st == LARVA]
            # agent 2: larva through iteration 17 (its threshold falls in closed hours), pupa after 18
            if t <= 18:
                assert lv["aux"].view(np.float32)[-1] == acc[t] and lv["hoursInState"][-1] == 2 + t
            else:  # it moved in ahead of the pupa that stayed (lower old index)
                pu = a[st == PUPA]
                assert len(pu) == 2 and pu["hoursInState"][0] == t - 19 and pu["ageInHours"][0] == 40 + t
                assert pu["aux"].view(np.float32)[0] == acc[19]
                assert pu["hoursInState"][1] == t - 3 and pu["ageInHours"][1] == 30 + t

    c = Case("larva_pupate", groups, 20, 1, steady_temp(20), env(), check)
    return c


def case_adult_transitions():
    groups = {PUPA: [dict(ageInHours=300, hoursInState=40, delay=40.5, isFemale=1),
                     dict(ageInHours=300, hoursInState=10, delay=100.0)],
              IMM: [dict(ageInHours=10, hoursInState=20, delay=20.0, isFemale=1)],
              MS: [dict(ageInHours=50, hoursInState=5, delay=1.0, isFemale=1),
                   dict(ageInHours=50, hoursInState=5, delay=1.0, isFemale=0)]}

    def check(rows, snaps):
        imm = immature_delay(27.3)
        lay1, a1 = snaps[1]
        st1 = states_of(lay1)
        im = a1[st1 == IMM]  # the emerged pupa
        assert len(im) == 1 and im["ageInHours"][0] == 0 and im["hoursInState"][0] == 0
        assert im["delay"][0] == imm and im["isFemale"][0] == 1
        ms = a1[st1 == MS]   # the matured immature moved in ahead of the two that stayed
        assert len(ms) == 3 and ms["delay"][0] == 1.0 and ms["hoursInState"][0] == 0
        assert list(ms["hoursInState"][1:]) == [6, 6]
        for t in range(2, 19):  # mate seeking leaves only at hour 18
            lay, a = snaps[t]
            assert (states_of(lay) == MS).sum() == 3 and (states_of(lay) == BMS).sum() == 0
        lay, a = snaps[19]
        st = states_of(lay)
        bms, ms = a[st == BMS], a[st == MS]
        assert len(bms) == 2 and (bms["isFemale"] == 1).all() and (bms["hoursInState"] == 0).all()
        assert (bms["delay"] == F32(CONST["species"]["bloodMealSeekingDelay"]["value"])).all()
        assert len(ms) == 1 and ms["isFemale"][0] == 0 and ms["hoursInState"][0] == 5 + 19
        # the emerged adult matures once hoursInState reaches the immature delay
        im = a[st == IMM]
        assert len(im) == 1 and im["hoursInState"][0] == 18 and im["ageInHours"][0] == 18

    return Case("adult_transitions", groups, 19, 1, steady_temp(19), env(), check)


def bms_group(n):
    """Blood-meal seeking females; those with i % 4 == 1 cross csd = 1 this
    hour (infectious if they bite), those with i % 4 == 3 stay just below."""
    return [dict(ageInHours=100, hoursInState=2, delay=0.0, isFemale=1, cycleLength=3,
                 csd=(0.995 if i % 2 else 0.5), humanBloodmealCount=(2 if i % 4 < 2 else 0)) for i in range(n)] + \
        [dict(ageInHours=100, hoursInState=2, delay=0.0, isFemale=1, cycleLength=3, csd=1.0, humanBloodmealCount=0)]


def case_bms_bite():
    n = 8
    groups = {BMS: bms_group(n)}
    lay0, _ = population(groups)
    g0 = gids_of(lay0)
    inf = (np.arange(n + 1) % 4 == 1) | (np.arange(n + 1) == n)  # at or past csd = 1 after this hour
    for seed in range(1, 5000):
        u = draws(seed, 0, 1, g0)
        bs = on_a_draw(u)
        bit = u <= bs
        if bit[n] and (bit & inf & (np.arange(n + 1) < n)).any() and (u < 1).all() and \
                (draws(seed, 0, 0, g0) > 0).all():
            break

    def check(rows, snaps):
        _, a0 = snaps[0]
        u = draws(seed, 0, 1, g0)
        bit = u <= bs
        inc = csd_increment(27.3)
        csd = np.array([F32(np.float64(x) + (1.0 * inc if h > 0 else 0.0)) for x, h in
                        zip(a0["csd"], a0["humanBloodmealCount"])], dtype=np.float32)
        lay1, a1 = snaps[1]
        st1 = states_of(lay1)
        bmd, bms = a1[st1 == BMD], a1[st1 == BMS]
        assert len(bmd) == bit.sum() and len(bms) == (~bit).sum()
        assert (bmd["humanBloodmealCount"] == a0["humanBloodmealCount"][bit] + 1).all()  # BITE_HUMAN 1: every bite is human
        assert (bmd["hoursInState"] == 0).all() and (bmd["delay"] == bmd_delay(27.3)).all()
        assert (bmd["cycleLength"] == 4).all() and (bms["cycleLength"] == 4).all()
        assert np.array_equal(bmd["csd"], csd[bit]) and np.array_equal(bms["csd"], csd[~bit])
        assert (bms["hoursInState"] == 3).all()
        assert rows[1]["bites"] == bit.sum() and bit[n] and not bit.all()  # the agent on the threshold bit
        assert rows[1]["infectBites"] == (bit & (csd >= 1.0)).sum()
        assert 1 < rows[1]["infectBites"] < rows[1]["bites"]  # csd exactly 1 counts

    return Case("bms_bite", groups, 1, seed, steady_temp(1), env(bloodmeal=float(bs)), check)


def case_itn():
    n = 8
    groups = {BMS: bms_group(n)}
    g0 = gids_of(population(groups)[0])
    seed = pick_seed(g0, 2, lambda u: u <= F32(0.6))
    itn = on_a_draw(draws(seed, 0, 2, g0))  # times BITE_HUMAN and BITE_INDOOR, both 1

    def check(rows, snaps):
        killed = draws(seed, 0, 2, g0) <= itn
        assert killed.any() and not killed.all()
        lay1, a1 = snaps[1]
        st1 = states_of(lay1)
        assert (st1 == BMD).sum() == (~killed).sum() and (st1 == BMS).sum() == 0
        assert rows[1]["bites"] == (~killed).sum()

    return Case("itn", groups, 1, seed, steady_temp(1), env(bloodmeal=1.0, itn=float(itn)), check)


def case_bmd_gravid():
    groups = {BMD: [dict(ageInHours=120, hoursInState=3, delay=2.0, isFemale=1, availableEggs=0, numEggBatches=2,
                         aux=7, cycleLength=5),
                    dict(ageInHours=120, hoursInState=3, delay=2.0, isFemale=1, availableEggs=5, numEggBatches=1,
                         aux=9, cycleLength=5),
                    dict(ageInHours=120, hoursInState=0, delay=50.0, isFemale=1, availableEggs=0, numEggBatches=0,
                         aux=4, cycleLength=5)]}
    g0 = gids_of(population(groups)[0])

    def check(rows, snaps):
        lay1, a1 = snaps[1]
        st1 = states_of(lay1)
        gr, bmd = a1[st1 == GRAVID], a1[st1 == BMD]
        assert len(gr) == 2 and len(bmd) == 1
        assert (gr["aux"] == 0).all() and (gr["hoursInState"] == 0).all() and (gr["cycleLength"] == 6).all()
        # the female without eggs made a batch: generateEggs with seeds + 1 (u0 = stream 1, u1 = stream 2)
        u1, u2 = draws(c.seed, 0, 1, g0[:1])[0], draws(c.seed, 0, 2, g0[:1])[0]
        e64 = box_muller64(u1, u2, CONST["generateEggs"]["mean"], CONST["generateEggs"]["sd"]) + 0.5
        decay = F32(np.float64(F32(CONST["species"]["eggBatchDecay"]["value"])) ** 2)  # pow(.8f, 2.0f)
        ok = set()
        for e0 in {int(np.floor(e64 - 1e-3)), int(np.floor(e64 + 1e-3))}:  # either side of a rounding edge
            ok.add(int(np.floor(F32(F32(max(e0, 0)) * decay) + F32(0.5))))
        assert int(gr["availableEggs"][0]) in ok, (gr["availableEggs"][0], ok)
        assert gr["numEggBatches"][0] == 3
        # the one that had eggs keeps them and its batch count
        assert gr["availableEggs"][1] == 5 and gr["numEggBatches"][1] == 1
        assert bmd["aux"][0] == 4 and bmd["cycleLength"][0] == 6 and bmd["hoursInState"][0] == 1

    c = Case("bmd_gravid", groups, 1, 1, steady_temp(1), env(), check)
    return c


def case_irs():
    groups = {BMD: [dict(ageInHours=120, hoursInState=0, delay=50.0, isFemale=1, cycleLength=5) for _ in range(8)]}
    g0 = gids_of(population(groups)[0])
    seed = pick_seed(g0, 2, lambda u: u <= F32(0.5))
    irs = on_a_draw(draws(seed, 0, 2, g0))  # times REST_INDOOR, 1

    def check(rows, snaps):
        killed = draws(seed, 0, 2, g0) <= irs
        lay1, a1 = snaps[1]
        bmd = a1[states_of(lay1) == BMD]
        assert len(bmd) == (~killed).sum() and (bmd["cycleLength"] == 6).all()

    return Case("irs", groups, 1, seed, steady_temp(1), env(irs=float(irs)), check)


def case_gravid_lay():
    cc = 20.0
    groups = {EGG: [dict(delay=1000.0) for _ in range(4)],
              LARVA: [dict(ageInHours=a, hoursInState=0, delay=1e9) for a in (30, 50, 100)],
              PUPA: [dict(ageInHours=300, hoursInState=0, delay=1000.0) for _ in range(2)],
              GRAVID: [dict(ageInHours=200, hoursInState=1, isFemale=1, availableEggs=100, aux=2, cycleLength=10,
                            numEggBatches=1) for _ in range(8)]}
    lay0, _ = population(groups)
    g_gr = np.arange(int(lay0[GRAVID][0]), int(lay0[GRAVID][1]))
    seed = pick_seed(g_gr, 1, lambda u: u < F32(0.5))
    lp = on_a_draw(draws(seed, 0, 1, g_gr))  # the agent on it does not lay: the test is strict

    def check(rows, snaps):
        lays = draws(seed, 0, 1, g_gr) < lp
        l1de = 30 // 24 + 50 // 24 + 100 // 24
        tb = l1de + 4 + 2  # killAgents.cl:42: L1DE + eggs.end + |pupae|
        nbm = F32(tb) / ((F32(2) + F32(1.0)) * F32(cc))
        pbm = max(F32(0.0), F32(1.0) - nbm)
        pe = int(np.ceil(F32(100) * pbm))
        assert pe == 79
        lay1, a1 = snaps[1]
        st1 = states_of(lay1)
        gr = a1[st1 == GRAVID]
        assert len(gr) == 8
        assert np.array_equal(gr["availableEggs"], np.where(lays, 100 - pe, 100))
        assert np.array_equal(gr["aux"], np.where(lays, 3, 2)) and (gr["cycleLength"] == 11).all()
        E = pe * int(lays.sum())
        eg = a1[st1 == EGG]
        assert len(eg) == E + 4
        # newborns: initAgents.cl:26-65 with this iteration's seeds 2 and 3 at indices 0..E-1
        nb = eg[:E]
        g = np.arange(E)
        p = draws(seed, 0, 2, g)
        fem = draws(seed, 0, 3, g).astype(np.float64) <= 0.5
        want = np.array([egg_incubation(27.3) + eggs_hatching(x) for x in p], dtype=np.float32)
        assert np.array_equal(nb["delay"], want)
        assert np.array_equal(nb["isFemale"], fem.astype(np.uint32))
        for f in ("ageInHours", "hoursInState", "csd", "aux", "availableEggs", "cycleLength",
                  "humanBloodmealCount", "numEggBatches"):
            assert (nb[f] == 0).all(), f
        old = eg[E:]
        assert (old["ageInHours"] == 1).all() and (old["delay"] == 1000.0).all()
        assert rows[1]["count"][0] == E + 4

    return Case("gravid_lay", groups, 1, seed, steady_temp(1), env(cc=cc), check,
                species={"lay_probability": float(lp)})


def case_gravid_cycle():
    groups = {GRAVID: [dict(ageInHours=200, hoursInState=1, isFemale=1, availableEggs=1, aux=0, cycleLength=7,
                            delay=3.0) for _ in range(6)]}
    g0 = gids_of(population(groups)[0])
    lp = CONST["species"]["layProbability"]["value"]
    seed = pick_seed(g0, 1, lambda u: u < F32(lp))

    def check(rows, snaps):
        lays = draws(seed, 0, 1, g0) < F32(lp)
        lay1, a1 = snaps[1]
        st1 = states_of(lay1)
        bms, gr = a1[st1 == BMS], a1[st1 == GRAVID]
        assert len(bms) == lays.sum() and len(gr) == (~lays).sum()
        assert (bms["cycleLength"] == 0).all() and (bms["hoursInState"] == 0).all()
        assert (bms["delay"] == F32(CONST["species"]["bloodMealSeekingDelay"]["value"])).all()
        assert (bms["aux"] == 1).all() and (bms["availableEggs"] == 0).all()
        assert (gr["cycleLength"] == 8).all() and (gr["aux"] == 0).all() and (gr["availableEggs"] == 1).all()
        assert rows[1]["numCycles"] == lays.sum() and rows[1]["sumCycles"] == 8 * lays.sum()
        assert rows[1]["count"][0] == lays.sum()  # one egg each

    return Case("gravid_cycle", groups, 1, seed, steady_temp(1), env(cc=100.0), check)


def case_ovitrap():
    groups = {GRAVID: [dict(ageInHours=200, hoursInState=1, isFemale=1, availableEggs=50, cycleLength=4)
                       for _ in range(12)]}
    g0 = gids_of(population(groups)[0])
    lp = CONST["species"]["layProbability"]["value"]
    seed, ov = None, None
    for s in range(1, 5000):  # some layers trapped, one of them on the threshold, some not
        lays = draws(s, 0, 1, g0) < F32(lp)
        if lays.sum() < 3 or not (draws(s, 0, 0, g0) > 0).all():
            continue
        u2 = draws(s, 0, 2, g0)
        ov = on_a_draw(u2[lays])
        trap = u2 <= ov
        if (lays & trap).any() and (lays & ~trap).any():
            seed = s
            break
    assert seed is not None

    def check(rows, snaps):
        lays = draws(seed, 0, 1, g0) < F32(lp)
        trap = lays & (draws(seed, 0, 2, g0) <= F32(ov))
        lay1, a1 = snaps[1]
        st1 = states_of(lay1)
        # trapped layers die, and their cycle still counts (updateAgents.cl:236-264)
        assert len(a1) - rows[1]["count"][0] == 12 - trap.sum()
        assert (st1 == BMS).sum() == (lays & ~trap).sum() and (st1 == GRAVID).sum() == (~lays).sum()
        assert rows[1]["numCycles"] == lays.sum() and rows[1]["sumCycles"] == 5 * lays.sum()
        assert rows[1]["count"][0] == 50 * (lays & ~trap).sum()

    return Case("ovitrap", groups, 1, seed, steady_temp(1), env(cc=1e9, ovitrap=float(ov)), check)


def case_csd():
    """Sporogonic development in binary64 then rounded (updateAgents.cl:69-70)
    over a varying temperature, including T = 16 and below; males never
    leave mate seeking, so they stay put for the whole run."""
    rng = np.random.default_rng(5)
    steps = 60
    temp = rng.uniform(10.0, 40.0, steps + 1).astype(np.float32)
    temp[[2, 5, 9]] = (16.0, 14.0, 16.5)
    start = [0.1, 0.001, 0.37, 2.0, 0.999, 17.1]
    groups = {MS: [dict(ageInHours=80, hoursInState=3, delay=1.0, isFemale=0, humanBloodmealCount=3, csd=c)
                   for c in start] +
              [dict(ageInHours=80, hoursInState=3, delay=1e9, isFemale=1, humanBloodmealCount=0, csd=0.3)]}

    def fold(binary64):
        c = np.float32(start)
        out = []
        for t in range(steps):
            cond = 1.0 if F32(temp[t]) > F32(16.0) else 0.0
            inc = cond * csd_increment(temp[t])
            c = (c.astype(np.float64) + inc).astype(np.float32) if binary64 else c + F32(inc)
            out.append(c.copy())
        return out

    want = fold(True)
    assert any((a != b).any() for a, b in zip(want, fold(False)))  # binary32 accumulation would differ

    def check(rows, snaps):
        for t in range(steps):
            lay, a = snaps[t + 1]
            assert np.array_equal(a["csd"][:len(start)], want[t]), t
            assert a["csd"][len(start)] == F32(0.3)

    return Case("csd", groups, steps, 1, temp, env(), check)


def case_delays_by_temperature():
    """The state-entry delays at seven temperatures, one transition of each
    kind per iteration in hours 0-6: pupaDelay in binary64 (its unsuffixed
    literals), immatureDelay and bloodMealDigestingDelay in binary32."""
    temp = np.float32([17.3, 21.7, 24.9, 28.1, 31.4, 35.6, 39.2, 30.0])
    acc, dly = F32(0.0), []
    for t in temp[:7]:
        acc = acc + dev_rate(t)
        dly.append(acc)  # larva j reaches its delay on the (j + 1)-th hour
    groups = {LARVA: [dict(ageInHours=60, hoursInState=0, delay=d) for d in dly],
              PUPA: [dict(ageInHours=300, hoursInState=0, delay=float(j + 1)) for j in range(7)],
              BMS: [dict(ageInHours=90, hoursInState=0, delay=float(j + 1), isFemale=1) for j in range(7)]}
    # binary32 evaluation of pupaDelay would differ at some of these temperatures
    assert any(pupa_delay(t) != F32(F32(-0.923) * t) + F32(60.923) for t in temp[:7])

    def check(rows, snaps):
        for j in range(7):
            lay, a = snaps[j + 1]
            st = states_of(lay)
            pu, im, bmd = a[st == PUPA], a[st == IMM], a[st == BMD]
            # newest arrivals first: each range is moved-in run, then stayers
            assert pu["delay"][0] == pupa_delay(temp[j]) and pu["hoursInState"][0] == 0
            assert im["delay"][0] == immature_delay(temp[j]) and im["ageInHours"][0] == 0
            assert bmd["delay"][0] == bmd_delay(temp[j]) and bmd["hoursInState"][0] == 0
            assert len(bmd) == j + 1 and rows[j + 1]["bites"] == 1

    return Case("delays_by_temperature", groups, 7, 1, temp, env(bloodmeal=1.0), check)


# seeds whose iteration-0 draws put one newborn exactly on a threshold of
# initAgents.cl: stream 3 on isFemale's 0.5 (<=), stream 2 on an
# eggsHatching band edge (p < threshold), found by searching with kat_rng
NEWBORN_EDGES = ((112010, 3, 67, 0.5), (2581, 2, 32, 0.94), (33515, 2, 26, 0.85), (39425, 2, 203, 0.5))


def case_newborn_edge(seed, stream, gid, value):
    n_eggs = 250
    groups = {GRAVID: [dict(ageInHours=200, hoursInState=1, isFemale=1, availableEggs=n_eggs, cycleLength=4)]}
    g0 = gids_of(population(groups)[0])

    def check(rows, snaps):
        assert draws(seed, 0, stream, [gid])[0] == F32(value)
        assert draws(seed, 0, 1, g0)[0] < F32(1.0) and draws(seed, 0, 0, g0)[0] > 0
        lay1, a1 = snaps[1]
        eg = a1[states_of(lay1) == EGG]
        assert len(eg) == n_eggs  # no biomass: all of them laid
        g = np.arange(n_eggs)
        p = draws(seed, 0, 2, g)
        fem = draws(seed, 0, 3, g).astype(np.float64) <= 0.5
        assert np.array_equal(eg["isFemale"], fem.astype(np.uint32))
        want = np.array([egg_incubation(27.3) + eggs_hatching(x) for x in p], dtype=np.float32)
        assert np.array_equal(eg["delay"], want)

    c = Case(f"newborn_edge_{seed}", groups, 1, seed, steady_temp(1, 27.3), env(cc=1e9), check,
             species={"lay_probability": 1.0})
    return c


def case_larvicide():
    groups = {LARVA: [dict(ageInHours=24, hoursInState=0, delay=1e9) for _ in range(6)] +
              [dict(ageInHours=60, hoursInState=3, delay=1e9) for _ in range(2)]}
    g0 = gids_of(population(groups)[0])
    seed = pick_seed(g0[:6], 2, lambda u: u <= F32(0.5))
    lv = on_a_draw(draws(seed, 0, 2, g0[:6]))

    def check(rows, snaps):
        killed = draws(seed, 0, 2, g0[:6]) <= lv
        lay1, a1 = snaps[1]
        n1 = (states_of(lay1) == LARVA).sum()
        assert n1 == 8 - killed.sum()
        lay2, a2 = snaps[2]  # only in their first hour as larvae
        assert (states_of(lay2) == LARVA).sum() == n1

    return Case("larvicide", groups, 2, seed, steady_temp(2), env(larvicide=float(lv)), check)


CASES = [case_egg_hatch, case_larva_pupate, case_adult_transitions, case_bms_bite, case_itn, case_bmd_gravid,
         case_irs, case_gravid_lay, case_gravid_cycle, case_ovitrap, case_csd, case_larvicide,
         case_delays_by_temperature] + [
    (lambda e: (lambda: case_newborn_edge(*e)))(e) for e in NEWBORN_EDGES]
for _f, _e in zip(CASES[-len(NEWBORN_EDGES):], NEWBORN_EDGES):
    _f.__name__ = f"case_newborn_edge_{_e[0]}"


def histogram_population():
    """Adults with ages at the bin edges of ageHistogram.cl:30 (days = hours
    / 24 in binary32, capped at bin 99), plus non-adults that must not count."""
    ages = [0, 23, 24, 25, 47, 48, 1000, 2376, 2399, 2400, 5000, 60000]
    return {EGG: [dict(delay=1000.0)], LARVA: [dict(ageInHours=500, delay=1e9)],
            IMM: [dict(ageInHours=a, delay=1e9) for a in ages[:4]],
            MS: [dict(ageInHours=a, delay=1e9, isFemale=0) for a in ages[4:8]],
            GRAVID: [dict(ageInHours=a, hoursInState=1, isFemale=1, availableEggs=1, aux=0) for a in ages[8:]]}


def histogram_of(layout, agents):
    """numpy.bincount of min(floor(age / 24), 99) over the adults."""
    adult = states_of(layout) >= IMM
    days = np.floor(agents["ageInHours"][adult] / F32(24.0)).astype(np.int64)
    return np.bincount(np.minimum(days, 99), minlength=100).astype(np.uint32)
