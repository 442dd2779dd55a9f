"""The autotuner's search (sampo_amd/batch.py) on the CPU with synthetic
timing models: coordinate descent finds the best value per size bucket of a
separable model, respects the limit on distinct tile shapes, and the
tuning's choices per initial size round-trip through its file."""
import numpy as np

from sampo_amd import batch as B


def test_descent_finds_each_buckets_best_of_a_separable_model():
    rng = np.random.default_rng(1)
    cands = list(B.CHUNK_CANDIDATES)
    # rate = sum over buckets of a per-bucket score with one best candidate each
    score = {b: dict(zip(cands, rng.permutation(len(cands)) + 1.0)) for b in range(6)}
    best = {b: max(cands, key=lambda c: score[b][c]) for b in range(6)}
    calls = []

    def measure(st):
        calls.append(dict(st))
        return sum(score[b][st[b]] for b in st)

    st, rate, table = B.coordinate_descent(list(range(6)), cands, {b: B.DEFAULT_CHUNK for b in range(6)}, measure)
    assert st == best and rate == sum(score[b][best[b]] for b in range(6))
    # the start, then every other candidate of each bucket (the start value again once it was left)
    assert 1 + 6 * (len(cands) - 1) <= len(calls) <= 1 + 6 * len(cands) and len(table) == len(calls)


def test_descent_keeps_the_start_when_nothing_is_faster():
    st, rate, _ = B.coordinate_descent([0, 1], [1, 2, 3], {0: 2, 1: 2}, lambda s: -abs(s[0] - 2) - abs(s[1] - 2))
    assert st == {0: 2, 1: 2} and rate == 0


def test_tile_shapes_beyond_the_limit_are_not_tried():
    tiles = [tuple(t) for t in B.TILE_CANDIDATES]
    default = tiles[0]
    # a model that would like a different shape for every bucket
    want = {b: tiles[b % len(tiles)] for b in range(6)}
    seen = []

    def measure(st):
        seen.append(set(st.values()) | {default})
        return sum(1.0 for b in st if st[b] == want[b])

    st, _, _ = B.coordinate_descent(list(range(6)), tiles, {b: default for b in range(6)}, measure,
                                    allowed=lambda t: len(set(t.values()) | {default}) <= B.MAX_TILE_SHAPES)
    assert all(len(s) <= B.MAX_TILE_SHAPES for s in seen)
    assert len(set(st.values()) | {default}) <= B.MAX_TILE_SHAPES
    assert sum(st[b] == want[b] for b in range(6)) >= B.MAX_TILE_SHAPES - 1


def test_tuning_per_size_and_file_round_trip(tmp_path):
    t = B.Tuning(tile=(768, 1, 6), chunk={0: 4096, 3: 65536}, tiles={0: (256, 1, 8), 5: (512, 2, 0)})
    assert t.chunk_for(1500) == 4096 and t.tile_for(1500) == (256, 1, 8)
    assert t.chunk_for(20_000) == B.DEFAULT_CHUNK and t.tile_for(20_000) == (768, 1, 6)
    assert t.chunk_for(100_000) == 65536
    assert B.bucket_of(3e6) == 5 and t.tile_for(3e6) == (512, 2, 0)
    p = tmp_path / "t.json"
    B.save_tuning(t, p)
    u = B.load_tuning(p)
    assert u.tile == t.tile and u.chunk == t.chunk and u.tiles == t.tiles
