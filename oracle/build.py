"""Build recipe for the oracle's shared library (test infrastructure only).

Compiles sampo_oracle.c with the canonical numeric flags into
oracle/liboracle.so.  The reference itself cannot be compiled here: its hot
path is OpenCL C that needs an OpenCL runtime and device, so there is no
oracle/_ref build.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math", "-o", LIB, os.path.join(HERE, "sampo_oracle.c"), "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
