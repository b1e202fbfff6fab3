"""The C-ABI libraries load without a GPU and export every entry point
their headers declare (include/rigc.h -> librigc.so,
rigid-ipc_amd/host/rigs.h -> librigc_sim.so). No compute call is made."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PAIRS = [
    (os.path.join(ROOT, "include", "rigc.h"), os.path.join(ROOT, "rigid-ipc_amd", "librigc.so"), "rigc_"),
    (os.path.join(ROOT, "rigid-ipc_amd", "host", "rigs.h"), os.path.join(ROOT, "rigid-ipc_amd", "librigc_sim.so"),
     "rigs_"),
]


def declared(header, prefix):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+(" + prefix + r"[a-z0-9_]+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib,prefix", PAIRS)
def test_library_exports_declared_symbols(header, lib, prefix):
    if not os.path.exists(lib):
        pytest.fail(f"{lib} is not built (run __graft_entry__.build())")
    names = declared(header, prefix)
    assert len(names) >= 6, names
    L = ctypes.CDLL(lib)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_rigc_header_declares_the_path():
    names = declared(PAIRS[0][0], "rigc_")
    for n in ("rigc_broad_phase", "rigc_narrow_phase", "rigc_compute_earliest_toi", "rigc_construct_constraint_set",
              "rigc_compute_barrier_term", "rigc_compute_friction_term", "rigc_update_friction_constraints",
              "rigc_compute_minimum_distance", "rigc_detect_intersections"):
        assert n in names
