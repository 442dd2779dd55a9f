"""Compile the gfx950 libraries of the package in-tree.

libsampo_engine.so  fused, batched agent-step engine (csrc/sampo_engine.hip)

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
    "libsampo_engine.so": ["sampo_engine.hip"],
}


def build(verbose=False):
    out = []
    for lib, srcs in TARGETS.items():
        dst = os.path.join(HERE, lib)
        cmd = ["hipcc", *HIP_FLAGS, "-o", dst, *[os.path.join(CSRC, s) for s in srcs]]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                         if "packed-fp32-ops" not in l and l.strip())
        if r.returncode != 0:
            sys.stderr.write(msgs + "\n")
            raise RuntimeError(f"hipcc failed for {lib}")
        if verbose and msgs:
            print(msgs)
        out.append(dst)
    return out


if __name__ == "__main__":
    for p in build(verbose=True):
        print(p)
