"""Launch parameters change nothing but speed (MI355X).

Instances admitted under different tile shapes are advanced by one pass
launch per shape; chunk sizes differ per size bucket.  A mixed set run with
four tile shapes and per-bucket chunks must give every instance the rows it
gets with the default launch parameters, and a lone instance of a second
shape (the one-slot planning path) the rows of its run alone.
"""
import numpy as np
import pytest

import sampo_amd as S
from sampo_amd import batch as B
from sampo_amd import workloads as W

pytestmark = pytest.mark.gpu

TUNED = B.Tuning(tile=(768, 1, 6),
                 chunk={0: 4096, 1: 65536, 2: 16384, 3: 262144, 4: 4096, 5: 65536},
                 tiles={0: (256, 1, 8), 1: (512, 2, 0), 2: (256, 2, 6), 3: (512, 2, 0), 4: (256, 1, 8)})


def run_set(specs, tuning, days=2):
    with S.Engine(max_slots=len(specs)) as e:
        slots = B.admit_specs(e, specs, tuning, phys_factor=1.25)
        for _ in range(days):
            e.step_days(1)
            e.service(1.25)
        st, step, _ = e.poll(len(specs))
        assert (st == S.RUNNING).all() and (step == 24 * days).all()
        return [e.rows(s, 0, 24 * days + 1) for s in slots]


def test_tuned_launches_equal_default_launches():
    specs = W.sweep(count=4096)[:64]
    buckets = {B.bucket_of(sp.n0) for sp in specs}
    assert len(buckets) >= 5 and len({TUNED.tile_for(sp.n0) for sp in specs}) == 4
    print(f"{len(specs)} scenarios, {sum(sp.n0 for sp in specs) / 1e6:.1f}M agents, buckets {sorted(buckets)}", flush=True)
    tuned = run_set(specs, TUNED)
    default = run_set(specs, B.Tuning())
    for i, (a, b) in enumerate(zip(tuned, default)):
        assert np.array_equal(a, b), f"scenario {i} (n0 {specs[i].n0}) differs with tuned launch parameters"


def test_lone_instance_of_a_second_shape():
    specs = W.sweep(count=4096)[:3]
    kws = [kw for _, kw in W.materialise_all(specs)]
    with S.Engine(max_slots=2) as e:
        e.set_tile(768, 1, 6)
        a = e.admit(**kws[0])          # slot 0, shape 768x1x6
        e.set_tile(256, 1, 8)
        b = e.admit(**kws[1])          # slot 1, a second shape
        e.release(a)
        c = e.admit(**kws[2])          # slot 0 again, second shape
        e.release(b)                   # one slot left: planned by its prep alone
        assert c == 0
        e.step_days(2)
        got = e.rows(c, 0, 49)
    with S.Engine(max_slots=1) as e:
        s = e.admit(**kws[2])
        e.step_days(2)
        alone = e.rows(s, 0, 49)
    assert np.array_equal(got, alone)
