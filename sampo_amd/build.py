"""Compile the gfx950 libraries of the package in-tree.

libsampo_engine.so  fused, batched agent-step engine (csrc/sampo_engine.hip)
libsampo_icl.so     the reference's icl_* runtime API, scan API and kernel
                    contract over HIP, plus the reference driver loop and its
                    text input/output (csrc/sampo_icl.hip, csrc/sampo_abms.c,
                    csrc/sampo_report.c)
grid_wait_probe     a measurement program: a grid-wide wait inside one launch
                    against a kernel boundary in a graph (csrc/grid_wait_probe.hip)

Every HIP source is compiled with -ffp-contract=off and IEEE division/sqrt so
that device arithmetic follows the canonical numeric contract bit for bit,
and with packed fp32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ARCH = os.environ.get("SAMPO_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

TARGETS = {
    "libsampo_engine.so": ["sampo_engine.hip", "sampo_compact.hip", "sampo_species.c"],
    "libsampo_icl.so": ["sampo_icl.hip", "sampo_abms.c", "sampo_report.c", "sampo_species.c"],
}
PROGRAMS = {"grid_wait_probe": "grid_wait_probe.hip"}
C_FLAGS = ["-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        sys.stderr.write(msgs + "\n")
        raise RuntimeError(f"build failed: {what}")
    if verbose and msgs:
        print(msgs)


def build(verbose=False):
    """Compile every source to an object (gcc for host C, hipcc for HIP) and
    link each library with hipcc."""
    out = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for lib, srcs in TARGETS.items():
        objs = []
        for s in srcs:
            obj = os.path.join(bdir, s.rsplit(".", 1)[0] + ".o")
            src = os.path.join(CSRC, s)
            if s.endswith(".c"):
                _run(["gcc", *C_FLAGS, "-c", "-o", obj, src], s, verbose)
            else:
                _run(["hipcc", *[f for f in HIP_FLAGS if f != "-shared"], "-c", "-o", obj, src], s, verbose)
            objs.append(obj)
        dst = os.path.join(HERE, lib)
        _run(["hipcc", "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", dst, *objs], lib, verbose)
        out.append(dst)
    for exe, src in PROGRAMS.items():
        dst = os.path.join(HERE, exe)
        _run(["hipcc", *[f for f in HIP_FLAGS if f not in ("-shared", "-fPIC")], "-o", dst, os.path.join(CSRC, src)],
             exe, verbose)
        out.append(dst)
    return out


if __name__ == "__main__":
    for p in build(verbose=True):
        print(p)
