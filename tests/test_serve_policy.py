"""The serving loop's day policy (sampo_amd/serve.py, Server._choose), on
the CPU with a stand-in engine: which running requests advance and which
are held."""
import numpy as np

import sampo_amd as S
from sampo_amd import serve as V
from sampo_amd import workloads as W


class Recorder:
    def __init__(self):
        self.masks = []

    def hold(self, mask):
        self.masks.append(np.asarray(mask, bool).copy())


def specs_of(n0s, days):
    return [W.Spec(index=i, seed=i, n0=n, days=d) for i, (n, d) in enumerate(zip(n0s, days))]


def poll_of(n, rows):
    st = np.full(n, S.EMPTY, np.uint32)
    step = np.zeros(n, np.uint32)
    na = np.zeros(n, np.uint64)
    for s, (a, b, c) in rows.items():
        st[s], step[s], na[s] = a, b, c
    return st, step, na


def test_least_work_left_advances_until_min_active():
    # work left = agents now x iterations left
    specs = specs_of([1000, 1000, 1000, 1000], [10, 10, 10, 10])
    running = {0: 0, 1: 1, 2: 2, 3: 3}
    last = poll_of(4, {0: (S.RUNNING, 0, 500_000), 1: (S.RUNNING, 200, 100_000),
                       2: (S.RUNNING, 100, 300_000), 3: (S.HELD, 0, 900_000)})
    # left: 0 -> 500k*240, 1 -> 100k*40, 2 -> 300k*140, 3 -> 900k*240
    e = Recorder()
    V.Server(min_active=350_000)._choose(e, running, specs, last, set())
    assert e.masks[-1].tolist() == [True, False, False, True]  # 1 then 2 make 400k >= 350k
    V.Server(min_active=1)._choose(e, running, specs, last, set())
    assert e.masks[-1].tolist() == [True, False, True, True]   # only the least work left
    V.Server(min_active=10**9)._choose(e, running, specs, last, set())
    assert not e.masks[-1].any()                               # all of them


def test_fresh_requests_count_from_their_initial_population():
    specs = specs_of([50_000, 10], [30, 30])
    running = {0: 0, 4: 1}              # slot 4 was admitted after the last poll
    last = poll_of(2, {0: (S.RUNNING, 0, 80_000)})
    e = Recorder()
    V.Server(min_active=1)._choose(e, running, specs, last, {4})
    m = e.masks[-1]
    assert m.shape[0] == 5 and m.tolist() == [True, False, False, False, False]


def test_other_states_are_left_alone():
    specs = specs_of([1000, 1000, 1000], [10, 10, 10])
    running = {0: 0, 1: 1, 2: 2}
    last = poll_of(3, {0: (S.GROW, 0, 10), 1: (S.RUNNING, 0, 5_000), 2: (S.DONE, 240, 7)})
    e = Recorder()
    V.Server(min_active=1)._choose(e, running, specs, last, set())
    assert e.masks[-1].tolist() == [False, False, False]      # growth and finished: not held


def test_unknown_policy_is_refused():
    try:
        V.Server(policy="fifo")
    except ValueError:
        return
    raise AssertionError("policy 'fifo' accepted")
