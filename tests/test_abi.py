"""The C ABI as declared in include/*.h (CPU: loads the built libraries and
checks their exports; no device calls), and that the model constants of
include/properties.h are written once, in the species table.
"""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(ROOT, "sampo_amd", "csrc")
LIBS = {
    "sampo_engine.h": os.path.join(ROOT, "sampo_amd", "libsampo_engine.so"),
    "sampo_compact.h": os.path.join(ROOT, "sampo_amd", "libsampo_engine.so"),
    "sampo_icl.h": os.path.join(ROOT, "sampo_amd", "libsampo_icl.so"),
    "sampo_abms.h": os.path.join(ROOT, "sampo_amd", "libsampo_icl.so"),
    "sampo_report.h": os.path.join(ROOT, "sampo_amd", "libsampo_icl.so"),
}


def declared_functions(header):
    """Names of the functions a header declares: top-level statements with a
    parameter list that are not typedefs."""
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"\\\n", " ", text)                     # join continued lines
    text = re.sub(r"^\s*#.*$", " ", text, flags=re.M)       # drop preprocessor lines
    text = text.replace('extern "C" {', " ")
    names, depth, stmt = [], 0, []
    for ch in text:
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth = max(0, depth - 1)
        if ch == ";" and depth == 0:
            st = " ".join("".join(stmt).split())
            stmt = []
            if "(" in st and not st.startswith("typedef"):
                m = re.search(r"(\w+)\s*\($", st.split("(")[0] + "(")
                if m:
                    names.append(m.group(1))
        elif depth == 0 and ch != "}":
            stmt.append(ch)
    return names


@pytest.mark.parametrize("header", sorted(LIBS))
def test_library_exports_every_declared_function(header):
    path = LIBS[header]
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (python -m sampo_amd.build)")
    names = declared_functions(header)
    assert names, f"no declarations parsed from {header}"
    lib = ctypes.CDLL(path)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{os.path.basename(path)} lacks {missing}"
    # and the dynamic symbol table says they are C (unmangled) text symbols
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.split()[-2:-1] in (["T"], ["t"])}
    assert set(names) <= exported


def test_headers_declare_the_reference_boundary():
    """The icl surface the reference driver uses (include/lib_icl.h:100-159,
    include/scan.h:5-9) is declared by the drop-in header."""
    names = set(declared_functions("sampo_icl.h"))
    for n in ("icl_init_devices", "icl_get_num_devices", "icl_get_device", "icl_release_devices",
              "icl_print_device_short_info", "icl_create_buffer", "icl_read_buffer", "icl_read_buffer_offset",
              "icl_write_buffer", "icl_write_buffer_offset", "icl_copy_buffer", "icl_release_buffer",
              "icl_release_buffers", "icl_create_kernel", "icl_run_kernel", "icl_release_kernel",
              "icl_init_timer", "icl_start_timer", "icl_stop_timer", "icl_release_timer", "icl_create_event",
              "icl_release_event", "icl_release_events", "icl_profile_event", "icl_finish",
              "segmented_scan_init", "segmented_scan_release", "scan"):
        assert n in names, n


def test_species_constants_live_only_in_the_species_table():
    """include/properties.h's numbers are written once, in sampo_species.c
    (checked against the reference in test_species.py); every kernel reads
    them from a species table."""
    for f in os.listdir(CSRC):
        if f == "sampo_species.c" or not f.endswith((".hip", ".h", ".c")):
            continue
        src = open(os.path.join(CSRC, f)).read()
        for lit in ("0.000175044", "68.5714", "-0.923", "2.667", "77.231", "170.0f", "0.25f"):
            assert lit not in src, (f, lit)


