"""Text input and output of the driver (CPU): printStats, printPopulation,
histogram.txt and the temperature / environment readers, against the
format strings and input lines the reference holds
(tests/golden/report_formats.json, temp_sample.txt, environment_sample.txt,
made by tests/golden/make_report_fixtures.py)."""
import json
import os

import numpy as np

import sampo_amd as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FMT = json.load(open(os.path.join(GOLD, "report_formats.json")))


def row(counts, l1de, females, potinf, ncyc, scyc, bites, ibites):
    r = np.zeros(1, S.ROW_DTYPE)[0]
    r["count"][:] = counts
    r["l1de"], r["females"], r["potInfective"] = l1de, females, potinf
    r["numCycles"], r["sumCycles"], r["bites"], r["infectBites"], r["valid"] = ncyc, scyc, bites, ibites, 1
    return r


def reference_line(step, hours, temp, r):
    """The line printStats prints, from its own format strings and the Stats
    fields of calcStats (src/abms.c:287-302)."""
    c = [int(x) for x in r["count"]]
    mature = sum(c[3:])
    l1de = int(r["l1de"]) if c[1] > 0 else 0
    f = FMT["printStats"]["formats"]
    cyc = int(r["sumCycles"]) // int(r["numCycles"]) if r["numCycles"] > 0 else 0
    t = float(np.float32(hours) * np.float32(step))
    return (f[0] % (t, float(np.float32(temp))) +
            f[1] % (mature, c[3], c[4], c[5], c[6], c[7], int(r["females"]), int(r["potInfective"]),
                    mature - int(r["females"])) +
            f[2] % (c[0], c[1], c[2], l1de, l1de + c[0] + c[2]) +
            f[3] % (cyc, int(r["bites"]), int(r["infectBites"])))


def test_stats_line_matches_reference_format():
    rng = np.random.default_rng(3)
    cases = [row([32000, 0, 0, 0, 0, 0, 0, 0], 0, 0, 0, 0, 0, 0, 0),
             row([5, 0, 7, 8, 9, 10, 11, 12], 99, 20, 2, 4, 50, 7, 1),  # no larvae: L1DE reported as 0
             row([1, 2, 3, 0, 0, 0, 0, 0], 4, 0, 0, 0, 17, 0, 0)]
    for _ in range(20):
        c = rng.integers(0, 2_000_000, 8)
        cases.append(row(c, rng.integers(0, 10**6), rng.integers(0, c[3:].sum() + 1), rng.integers(0, 1000),
                         rng.integers(0, 500), rng.integers(0, 10**5), rng.integers(0, 10**4), rng.integers(0, 100)))
    for step, r in enumerate(cases):
        for hours, temp in ((1.0, 30.0), (1.0, 27.35), (0.5, -3.25)):
            assert S.format_stats(step * 37, temp, r, hours) == reference_line(step * 37, hours, temp, r)


def test_population_table_matches_reference_format():
    lay = np.array([[0, 10], [64, 70], [128, 128], [192, 200], [256, 300], [320, 321], [384, 390], [448, 460]])
    want = "".join(lab + FMT["printRange"]["format"] % (a, b, b - a)
                   for lab, (a, b) in zip(FMT["printPopulation"]["labels"], lay))
    assert S.format_population(lay) == want


def test_histogram_matches_reference_format():
    h = np.random.default_rng(1).integers(0, 10**6, 100)
    assert S.format_histogram(h) == "".join(FMT["histogram"]["format"] % int(x) for x in h)


def test_read_temperature_file():
    path = os.path.join(GOLD, "temp_sample.txt")
    want = np.float32([float(x) for x in open(path).read().split()])
    got = S.read_single_val_file(path, 1000)
    assert len(got) == len(want) == 48 and np.array_equal(got, want)
    assert len(S.read_single_val_file(path, 10)) == 10


def test_read_environment_file(tmp_path):
    path = os.path.join(GOLD, "environment_sample.txt")
    env = S.read_environment(path, 100, 160000.0)
    assert len(env) == 48
    assert np.all(env["bloodmealSuccess"] == np.float32(0.25)) and np.all(env["carryingCapacity"] == 160000.0)
    # interventions: each is coverage x effectiveness in binary32
    p = tmp_path / "env.txt"
    vals = np.random.default_rng(5).uniform(0, 1, (6, 9)).astype(np.float32)
    p.write_text("".join(", ".join(repr(float(v)) for v in line) + " \n" for line in vals))
    env = S.read_environment(str(p), 10, 5.0)
    assert len(env) == 6
    for f, k in (("IRSValue", 0), ("ITNValue", 2), ("larvacideValue", 4), ("oviTrapValue", 6)):
        assert np.array_equal(env[f], vals[:, k] * vals[:, k + 1]), f
    assert np.array_equal(env["bloodmealSuccess"], vals[:, 8])
