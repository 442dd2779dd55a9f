"""Serving: requests that arrive while others run share the engine; each
result must be the same alone and in the mix, bit for bit (GPU)."""
import numpy as np
import pytest

import sampo_amd as S
from sampo_amd import serve as V
from sampo_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def trace():
    rng = np.random.default_rng(77)
    specs = []
    for r in range(10):
        mult = rng.uniform(0.5, 1.5, size=4)
        specs.append(W.Spec(index=r, seed=5000 + r, n0=int(rng.integers(1500, 20000)), days=int(rng.integers(3, 9)),
                            t_mean=float(rng.uniform(18, 32)), t_amp=float(rng.uniform(2, 8)),
                            dmr=tuple(W.DMR[c] * float(mult[c]) for c in range(4)), cc_factor=0.0125,
                            arrival=float(rng.integers(0, 7))))
    return specs


@pytest.mark.parametrize("policy,min_active", [("srpt", 64_000_000), ("srpt", 1), ("all", 0)])
def test_each_request_is_the_same_alone_and_in_the_mix(policy, min_active):
    """min_active = 1: every day only the request with the least work left
    advances and the others are held."""
    specs = trace()
    # 4 slots for 10 requests: later arrivals queue, slots are reused, and
    # instances start with only their initial population's worth of slots
    srv = V.Server(max_slots=4, phys_factor=1.0, headroom=1.0, keep_rows=True, policy=policy, min_active=min_active)
    results, summary = srv.run(specs, clock="rounds")
    assert summary["completed"] == len(specs)
    assert any(r.admitted > r.arrival for r in results)  # some waited for a slot
    for sp, r in zip(specs, results):
        assert r.status == S.DONE and r.steps == sp.horizon
        alone = V.run_alone(sp)
        assert np.array_equal(alone, r.rows), f"request {sp.index} differs from its run alone"
        assert r.digest == V.rows_digest(alone)
    # and one of them against the oracle
    sp = specs[3]
    kw = W.materialise(sp)
    sim = O.Sim(temp=kw["temp"], env=kw["env"], capacity=sp.max_agents, seed=sp.seed, layout=kw["layout"],
                agents=kw["agents"], mode=1, dmr=sp.dmr)
    orows = sim.run(sp.horizon)
    got = results[3].rows
    assert np.array_equal(orows["count"], got["count"][:sp.horizon])
    for f in ("l1de", "females", "potInfective", "numCycles", "sumCycles", "bites", "infectBites"):
        assert np.array_equal(orows[f], got[f][:sp.horizon]), f


def test_engine_hold_pauses_and_resumes_an_instance():
    """A held instance does not advance while the other runs; resumed, it
    finishes with the rows of its run alone, and so does the other."""
    days = 4
    kws = [W.materialise(W.Spec(index=k, seed=900 + k, n0=3000 + 2000 * k, days=days, t_mean=26.0, t_amp=4.0,
                                dmr=W.DMR, cc_factor=0.0125, arrival=0.0)) for k in range(2)]
    alone = []
    for kw in kws:
        with S.Engine(max_slots=1) as e:
            s = e.admit(**kw)
            e.step_days(days)
            alone.append(e.rows(s, 0, 24 * days + 1))
    with S.Engine(max_slots=2) as e:
        a, b = e.admit(**kws[0]), e.admit(**kws[1])
        e.step_days(1)
        mask = np.zeros(2, bool)
        mask[b] = True
        e.hold(mask)
        e.step_days(2)
        st, step, _ = e.poll(2)
        assert st[b] == S.HELD and step[b] == 24 and step[a] == 72
        e.hold(np.zeros(2, bool))
        e.step_days(1)
        st, step, _ = e.poll(2)
        assert st[a] == S.DONE and st[b] == S.RUNNING and step[b] == 48
        e.step_days(2)
        st, step, _ = e.poll(2)
        assert st[a] == S.DONE and st[b] == S.DONE
        assert np.array_equal(e.rows(a, 0, 24 * days + 1), alone[0])
        assert np.array_equal(e.rows(b, 0, 24 * days + 1), alone[1])
