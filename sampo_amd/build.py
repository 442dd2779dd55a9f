"""Compile the gfx950 libraries of the package in-tree.

libsampo_engine.so  fused, batched agent-step engine (csrc/sampo_engine.hip)
libsampo_icl.so     the reference's icl_* runtime API, scan API and kernel
                    contract over HIP, plus the reference driver loop and its
                    text input/output (csrc/sampo_icl.hip, csrc/sampo_abms.c,
                    csrc/sampo_report.c)

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
    "libsampo_engine.so": ["sampo_engine.hip", "sampo_species.c"],
    "libsampo_icl.so": ["sampo_icl.hip", "sampo_abms.c", "sampo_report.c", "sampo_species.c"],
}
C_FLAGS = ["-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]


def build(verbose=False):
    out = []
    for lib, srcs in TARGETS.items():
        dst = os.path.join(HERE, lib)
        objs = []
        for s in srcs:
            if s.endswith(".c"):  # host C: gcc, canonical flags
                obj = os.path.join(HERE, "_build", s[:-2] + ".o")
                os.makedirs(os.path.dirname(obj), exist_ok=True)
                subprocess.run(["gcc", *C_FLAGS, "-c", "-o", obj, os.path.join(CSRC, s)], check=True)
                objs.append(obj)
        hip_srcs = [os.path.join(CSRC, s) for s in srcs if not s.endswith(".c")]
        cmd = ["hipcc", *HIP_FLAGS, "-o", dst, *hip_srcs]
        if objs:  # hipcc reads every input as HIP source: link host objects in a second step
            hobj = os.path.join(HERE, "_build", lib + ".hip.o")
            cmd = ["hipcc", *[f for f in HIP_FLAGS if f != "-shared"], "-c", "-o", hobj, *hip_srcs]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError(f"hipcc failed for {lib}")
            cmd = ["hipcc", "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", dst, hobj, *objs]
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
