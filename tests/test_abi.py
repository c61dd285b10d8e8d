"""The C-ABI library loads without a GPU and exports every function that
include/mcgpu.h declares (no compute call is made)."""
import ctypes as C
import os
import re

from mcsolver_amd import _lib as L

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "mcgpu.h")


def declared_functions():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mcg_\w+)\s*\(", txt)))


def test_header_declares_the_binding_exports():
    assert declared_functions() == sorted(L.EXPORTS)


def test_library_exports_every_declared_function():
    lib = C.CDLL(L.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing
    assert L.lib().mcg_version() >= 1


def test_struct_sizes_match_header_layout():
    # mcg_batch: 12 int32 (48 B) + 17 pointers, then the exact-Ising table
    # pointer, its row length (int32) and 1/unit (float)
    assert C.sizeof(L.Batch) == 48 + 17 * 8 + 8 + 4 + 4
    assert L.Batch.ithr.offset == 48 + 17 * 8
    assert C.sizeof(L.Plan) == 64
    # mcg_graph: 14 int32 (56 B) + 17 pointers, then pair_s2 and two int32
    assert C.sizeof(L.Graph) == 56 + 17 * 8 + 8 + 8
    assert L.Graph.pair_s2.offset == 56 + 17 * 8
    assert C.sizeof(L.Replica) == 40
