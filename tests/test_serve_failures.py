"""The serving loop's failure handling (sampo_amd/serve.py, Server.run), on
the CPU with a stand-in engine: an admission the device refuses rejects
that request only, and an instance whose growth keeps failing is ended
after max_deferred days while the others complete."""
import numpy as np
import pytest

import sampo_amd as S
from sampo_amd import serve as V
from sampo_amd import workloads as W


class FakeEngine:
    """Slots advance 24 iterations per day unless held; the request with
    seed `refuse` cannot be admitted; the one with seed `stuck` pauses for a
    growth after its first day and every growth fails."""

    def __init__(self, refuse, stuck, max_slots=16):
        self.refuse, self.stuck = refuse, stuck
        self.slots = {}
        self.max_slots = max_slots
        self.released = []

    def __call__(self, max_slots=16, device=0):
        return self

    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False

    def set_tile(self, *a):
        pass

    def set_chunk(self, *a):
        pass

    def admit(self, **kw):
        if kw["seed"] == self.refuse:
            raise S.SampoError("device allocation failed")
        slot = min(set(range(self.max_slots)) - set(self.slots))
        self.slots[slot] = dict(seed=kw["seed"], horizon=kw["horizon"], step=0, status=S.RUNNING, held=False)
        return slot

    def hold(self, mask):
        for s, x in self.slots.items():
            x["held"] = bool(mask[s]) if s < len(mask) else False

    def step_days(self, n):
        for x in self.slots.values():
            if x["status"] == S.RUNNING and not x["held"]:
                x["step"] += 24
                if x["seed"] == self.stuck and x["step"] >= 24:
                    x["status"] = S.GROW
                elif x["step"] >= x["horizon"]:
                    x["status"] = S.DONE

    def service(self, headroom):
        if any(x["status"] == S.GROW for x in self.slots.values()):
            raise S.SampoError("growing an instance failed")
        return 0

    def poll(self, n):
        st = np.full(n, S.EMPTY, np.uint32)
        step = np.zeros(n, np.uint32)
        na = np.zeros(n, np.uint64)
        for s, x in self.slots.items():
            if s < n:
                st[s], step[s], na[s] = x["status"], x["step"], 1000
        return st, step, na

    def rows(self, slot, first, n):
        r = np.zeros(n, S.ROW_DTYPE)
        r["count"][:, 0] = self.slots[slot]["seed"]
        return r

    def release(self, slot):
        self.released.append(self.slots.pop(slot)["seed"])


@pytest.fixture
def fake(monkeypatch):
    eng = FakeEngine(refuse=3, stuck=5)
    monkeypatch.setattr(S, "Engine", eng)
    monkeypatch.setattr(S, "footprint", lambda chunk=0, **kw: 1 << 20)
    monkeypatch.setattr(S, "device_memory", lambda device=0: (1 << 40, 1 << 40))
    monkeypatch.setattr(W, "materialise", lambda sp, phys_factor=0.0: dict(seed=sp.seed, horizon=sp.horizon))
    return eng


def test_refused_admission_and_failed_growth_end_only_their_requests(fake):
    specs = [W.Spec(index=i, seed=i, n0=1000, days=3 + i % 3, arrival=0.0) for i in range(8)]
    results, summary = V.Server(max_slots=16, policy="all", max_deferred=4).run(specs, clock="rounds")
    by = {r.index: r for r in results}
    assert by[3].status == V.REJECTED
    assert by[5].status == V.GROWTH_FAILED and by[5].steps == 24
    for i in set(range(8)) - {3, 5}:
        assert by[i].status == S.DONE and by[i].steps >= specs[i].horizon
    assert summary["completed"] == 6 and summary["rejected"] == 1 and summary["growth_failed"] == 1
    assert sorted(fake.released) == [0, 1, 2, 4, 5, 6, 7]
    # ended after max_deferred days paused, not before
    assert by[5].done - by[5].admitted >= 4
