"""Admission limits and failures of the engine (MI355X).

  - a run whose humanBloodmealCount or numEggBatches could outgrow its
    16/15-bit field within the horizon is refused with SAMPO_E_ARG;
  - admitting instances until the device is out of memory fails cleanly:
    the refused admission leaves its slot empty, and the instances already
    running go on exactly as they would alone.
"""
import numpy as np
import pytest

import sampo_amd as S

pytestmark = pytest.mark.gpu

STATS = ("l1de", "females", "potInfective", "numCycles", "sumCycles", "bites", "infectBites")


def ref_kw(n_eggs=2000, days=2, max_agents=200_000, **kw):
    steps = 24 * days
    return dict(mode=S.MODE_REFERENCE, seed=7, horizon=steps, temp=np.full(steps, 30.0, np.float32),
                env=S.make_env(cc=10_000.0), n_eggs=n_eggs, max_agents=max_agents, **kw)


def test_admission_refuses_counters_that_could_wrap():
    temp = S.gen_temperature(3, 4)
    lay, ag = S.gen_population(3, 1000, float(temp[0]))
    env = S.make_env(cc=5000.0)
    with S.Engine(max_slots=2) as e:
        e.admit(mode=S.MODE_CONFIG, seed=3, horizon=96, temp=temp, env=env, max_agents=8192, layout=lay, agents=ag)
        for field, near in (("humanBloodmealCount", 0xffff - 10), ("numEggBatches", 0x7fff - 10)):
            bad = ag.copy()
            bad[field][5] = near  # 96 iterations allow 32 more cycles: could wrap
            with pytest.raises(S.SampoError):
                e.admit(mode=S.MODE_CONFIG, seed=3, horizon=96, temp=temp, env=env, max_agents=8192, layout=lay,
                        agents=bad)
        st, _, _ = e.poll(2)
        assert st[0] == S.RUNNING and st[1] == S.EMPTY


def test_out_of_memory_admission_leaves_the_others_intact():
    free, total = S.device_memory()
    big = ref_kw(max_agents=1 << 30, phys_slots=1 << 30)
    need = S.footprint(**big)
    assert need > 50 * 2**30
    k = int(free // need) + 1  # one more than fits
    assert k <= 8, f"{k} instances of {need / 2**30:.0f} GiB would be needed to fill {free / 2**30:.0f} GiB"
    print(f"admitting up to {k} instances of {need / 2**30:.1f} GiB into {free / 2**30:.1f} GiB free", flush=True)
    with S.Engine(max_slots=k + 1) as e:
        small = e.admit(**ref_kw())
        slots, refused = [], None
        for _ in range(k):
            try:
                slots.append(e.admit(**big))
            except S.SampoError as err:
                refused = str(err)
                break
        assert refused is not None and "allocation" in refused, refused
        st, _, _ = e.poll(k + 1)
        assert st[small] == S.RUNNING and all(st[s] == S.RUNNING for s in slots)
        assert (st[len(slots) + 1:] == S.EMPTY).all()
        for s in slots:
            e.release(s)
        e.step_days(2)
        rows = e.rows(small, 0, 48)
    with S.Engine(max_slots=1) as e:
        s = e.admit(**ref_kw())
        e.step_days(2)
        alone = e.rows(s, 0, 48)
    assert np.array_equal(rows["count"], alone["count"])
    for f in STATS:
        assert np.array_equal(rows[f], alone[f]), f
