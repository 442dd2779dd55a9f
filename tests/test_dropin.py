"""The drop-in recipe, applied to the reference's own driver (CPU).

The recipe is the set of edits a maintainer makes so that the reference's
src/abms.c builds against libsampo_icl.so instead of its OpenCL device
layer: include/abms.h includes sampo_icl.h in place of lib_icl.h,
lib_icl_ext.h and host_types.h, and src/abms.c calls icl_finish and
icl_wait_for_events where it called clFinish and clWaitForEvents; kernel
creation stays as it is and the link line names the library.  This test
applies exactly those edits to a scratch
copy of the reference's src/ and include/ (nothing of it is kept), compiles
src/abms.c with the reference Makefile's flags, links it against the
library and runs the binary's -h path (which touches no device).  It needs
the reference tree and is skipped where it is absent.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
LIBDIR = os.path.join(ROOT, "sampo_amd")

pytestmark = pytest.mark.skipif(not os.path.isfile(os.path.join(REF, "src", "abms.c")),
                                reason="the reference tree is not present")


def apply_recipe(tree):
    """The recipe's two source edits (its other steps change nothing in the
    source: kernel creation stays, the link line is the command below)."""
    p = os.path.join(tree, "include", "abms.h")
    s = open(p).read()
    old = '#include "lib_icl.h"\n#include "lib_icl_ext.h"\n\n#include "host_types.h"\n'
    assert s.count(old) == 1, "include/abms.h:7-10 is not as the recipe expects"
    open(p, "w").write(s.replace(old, '#include "sampo_icl.h"\n'))
    p = os.path.join(tree, "src", "abms.c")
    s = open(p).read()
    n_finish = s.count("clFinish(dev->queue)")
    s = s.replace("clFinish(dev->queue)", "icl_finish(dev)")
    s, n_wait = re.subn(r"clWaitForEvents\((1u?), (\w+)->event\)", r"icl_wait_for_events(\1, \2)", s)
    assert n_finish >= 1 and n_wait == 2
    assert "clFinish" not in s and "clWaitForEvents" not in s
    open(p, "w").write(s)


@pytest.mark.parametrize("scan", ["parallel", "serial"])
def test_reference_driver_compiles_and_links_against_the_drop_in(tmp_path, scan):
    lib = os.path.join(LIBDIR, "libsampo_icl.so")
    assert os.path.exists(lib), "build the library first"
    tree = tmp_path / "sampo"
    for d in ("src", "include"):
        shutil.copytree(os.path.join(REF, d), tree / d)
    apply_recipe(str(tree))
    exe = tree / "abms"
    # the reference Makefile's CC and CFLAGS (Makefile:2-3), its -Iinclude, and the drop-in
    cmd = ["gcc", "-fms-extensions", "-Wall", "-g", "-O0", "-std=c99", str(tree / "src" / "abms.c"),
           f"-I{tree / 'include'}", f"-I{os.path.join(ROOT, 'include')}", f"-L{LIBDIR}", "-lsampo_icl",
           f"-Wl,-rpath,{LIBDIR}", "-lm", "-lrt", "-o", str(exe)]
    if scan == "serial":  # the CPU-device build: serialScan instead of the Bolt scan (src/abms.c:404-440)
        src = open(tree / "include" / "agent.h").read()
        assert src.count("#define DEVICE_TYPE ICL_GPU") == 1
        open(tree / "include" / "agent.h", "w").write(src.replace("#define DEVICE_TYPE ICL_GPU",
                                                                  "#define DEVICE_TYPE ICL_CPU"))
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "error" not in r.stderr
    out = subprocess.run([str(exe), "-h"], capture_output=True, text=True, timeout=60)
    assert "Usage:" in out.stdout
    # every icl_* and scan symbol the driver uses resolves in the library
    und = subprocess.run(["nm", "-u", str(exe)], capture_output=True, text=True, check=True).stdout.split()
    used = sorted({s for s in und if s.startswith("icl_") or s in ("scan", "segmented_scan_init",
                                                                   "segmented_scan_release")})
    exported = set(subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                                  check=True).stdout.split())
    assert used and all(s in exported for s in used), [s for s in used if s not in exported]
