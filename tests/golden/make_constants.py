"""Write tests/golden/constants.json: numeric model constants read as text
from the reference sources (include/properties.h, kernel/killAgents.cl,
src/abms.c).  The JSON is data -- numbers with the file:line they come from --
used to pin the oracle's transcription of those constants.

Run in the container that has the reference mounted:
    python tests/golden/make_constants.py /root/reference
"""
import json
import os
import re
import sys


def grab(path, pattern, flags=0):
    text = open(path).read()
    m = re.search(pattern, text, flags)
    if not m:
        raise SystemExit(f"pattern not found in {path}: {pattern}")
    line = text[:m.start()].count("\n") + 1
    return m, line


def hour_mask(expr):
    """The hours 0..23 at which a transition-time predicate holds, from its
    return expression: `true`, or terms `worldTime <op> <n>[u]` joined by ||."""
    expr = expr.strip()
    while expr.startswith("(") and expr.endswith(")"):
        expr = expr[1:-1].strip()
    if expr == "true":
        return (1 << 24) - 1
    ops = {"<=": lambda h, n: h <= n, ">=": lambda h, n: h >= n, "==": lambda h, n: h == n,
           "<": lambda h, n: h < n, ">": lambda h, n: h > n}
    mask = 0
    for term in expr.split("||"):
        m = re.fullmatch(r"\s*worldTime\s*(<=|>=|==|<|>)\s*(\d+)u?\s*", term)
        if not m:
            raise SystemExit(f"unrecognised transition predicate: {expr}")
        mask |= sum(1 << h for h in range(24) if ops[m.group(1)](h, int(m.group(2))))
    return mask


def species(props, update):
    """The members of the species table that the header fixes beyond the
    delay coefficients above: transition-hour windows, bite/rest
    probabilities, the larvaDevelopment clamp, constant delays, the egg-batch
    decay, and the egg-laying probability of updateAgents.cl."""
    out = {}
    text = open(props).read()
    preds = ["eggsTransitionTime", "larvaeTransitionTime", "pupaeTransitionTime", "immaturesTransitionTime",
             "mateSeekingsTransitionTime", "bloodMealSeekingsTransitionTime", "bloodMealDigestingsTransitionTime",
             "gravidsEggLayTime"]
    out["window"] = []
    for fn in preds:
        m, line = grab(props, r"bool " + fn + r"\(UINT worldTime\) \{\s*return\s*([^;]*);", re.S)
        out["window"].append({"fn": fn, "mask": hour_mask(m.group(1)), "source": f"include/properties.h:{line}"})
    for d in ("BITE_HUMAN_PROBABILITY", "BITE_INDOOR_PROBABILITY", "REST_INDOOR_PROBABILITY"):
        m, line = grab(props, r"#define " + d + r" ([0-9.]+)f")
        out[d] = {"value": float(m.group(1)), "source": f"include/properties.h:{line}"}
    m, line = grab(props, r"min\(max\(temperature, ([0-9.]+)f\), ([0-9.]+)f\)")
    m2, _ = grab(props, r"devRates\[\(\(int\)clampedTemp\) - (\d+)\]")
    out["larvaDevelopment"] = {"clamp": [float(m.group(1)), float(m.group(2))], "base": int(m2.group(1)),
                               "source": f"include/properties.h:{line}"}
    for fn in ("mateSeekingDelay", "bloodMealSeekingDelay"):
        m, line = grab(props, r"REAL " + fn + r"\(Temperature temperature\) \{\s*return ([0-9.]+)f;", re.S)
        out[fn] = {"value": float(m.group(1)), "source": f"include/properties.h:{line}"}
    m, line = grab(props, r"pow\(([0-9]*\.[0-9]+)f,")
    out["eggBatchDecay"] = {"value": float(m.group(1)), "source": f"include/properties.h:{line}"}
    m, line = grab(update, r"shouldLayEggs = HybridTaus\(&seeds\[1\]\) < ([0-9.]+)f")
    out["layProbability"] = {"value": float(m.group(1)), "source": f"kernel/updateAgents.cl:{line}"}
    return out


def main(ref):
    props = os.path.join(ref, "include", "properties.h")
    out = {}
    m, line = grab(props, r"devRates\[\]\s*=\s*\{([^}]*)\}", re.S)
    nums = [float(x) for x in re.findall(r"[0-9.eE+-]+", re.sub(r"//.*", "", m.group(1)))]
    out["devRates"] = {"values": nums, "source": f"include/properties.h:{line}"}
    m, line = grab(props, r"REAL eggsHatching\(REAL probability\) \{(.*?)return", re.S)
    body = m.group(1)
    thr = [float(x.rstrip("f")) for x in re.findall(r"probability < ([0-9.]+f?)", body)]
    ms = [float(x.rstrip("f")) for x in re.findall(r"m\s*=\s*(-?[0-9.]+f?);", body)]
    bs = [float(x.rstrip("f")) for x in re.findall(r"b\s*=\s*(-?[0-9.]+f?);", body)]
    out["eggsHatching"] = {"thresholds": thr, "m": ms[1:], "b": bs[1:], "source": f"include/properties.h:{line}"}

    def linear(name):
        mm, ln = grab(props, name + r"\(Temperature temperature\) \{\s*return\s*([^;]*);", re.S)
        return {"expr": mm.group(1).strip(), "source": f"include/properties.h:{ln}"}

    for fn in ("eggsIncubation", "pupaDelay", "immatureDelay", "bloodMealDigestingDelay"):
        out[fn] = linear("REAL " + fn)
    m, line = grab(props, r"REAL eggBatchSize = ([0-9.]+)f;\s*REAL stdDev = ([0-9.]+)f;", re.S)
    out["generateEggs"] = {"mean": float(m.group(1)), "sd": float(m.group(2)), "source": f"include/properties.h:{line}"}
    m, line = grab(props, r"REAL mean = ([0-9.]+)f;\s*REAL stdDev = ([0-9.]+)f;", re.S)
    out["larvaDelay"] = {"mean": float(m.group(1)), "sd": float(m.group(2)), "source": f"include/properties.h:{line}"}
    kill = os.path.join(ref, "kernel", "killAgents.cl")
    m, line = grab(kill, r"a = ([0-9.]+)f;\s*B = ([0-9.]+)f;\s*s = ([0-9.]+)f;", re.S)
    out["adultMortality"] = {"a": float(m.group(1)), "B": float(m.group(2)), "s": float(m.group(3)),
                             "source": f"kernel/killAgents.cl:{line}"}
    gr = os.path.join(ref, "include", "gpuRand.h")
    m, line = grab(gr, r"return ([0-9.e-]+)f \*")
    out["rngScale"] = {"literal": m.group(1), "source": f"include/gpuRand.h:{line}"}
    steps = re.findall(r"TausStep\(z(\d), (\d+), (\d+), (\d+), (\d+)UL\)", open(gr).read())
    out["taus"] = [[int(a) for a in s] for s in steps]
    lcg = re.search(r"LCGStep\(z4, (\d+), (\d+)UL\)", open(gr).read())
    out["lcg"] = [int(lcg.group(1)), int(lcg.group(2))]
    # how HybridTaus forms each component's state from the seed and the
    # global id: z_i = seeds->z_i [+ offset] [+ gid] or seeds->z_i * gid
    gtext = open(gr).read()
    state = []
    for m in re.finditer(r"UINT z(\d) = seeds->z\d((?: [+*] (?:\d+u|gid))*);", gtext):
        terms = re.findall(r"([+*]) (\d+u|gid)", m.group(2))
        state.append({"z": int(m.group(1)),
                      "add": sum(int(v.rstrip("u")) for op, v in terms if op == "+" and v != "gid"),
                      "add_gid": any(op == "+" and v == "gid" for op, v in terms),
                      "mul_gid": any(op == "*" and v == "gid" for op, v in terms),
                      "source": f"include/gpuRand.h:{gtext[:m.start()].count(chr(10)) + 1}"})
    out["hybridState"] = state
    abms = os.path.join(ref, "src", "abms.c")
    txt = open(abms).read()
    out["abms"] = {"FACTOR": int(re.search(r"#define FACTOR (\d+)", txt).group(1)),
                   "initialAgentCount": re.search(r"UINT initialAgentCount = ([^;]*);", txt).group(1),
                   "carryingCapacity": re.search(r"REAL carryingCapacity = ([^;]*);", txt).group(1),
                   "maxSteps": int(re.search(r"long long maxSteps = (\d+);", txt).group(1))}
    out["species"] = species(props, os.path.join(ref, "kernel", "updateAgents.cl"))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "constants.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print(dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
