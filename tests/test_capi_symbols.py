"""The C-ABI library loads and exports every function include/piqmc_hip.h
declares (no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import piqmc_capi as capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "piqmc_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pq_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "pq_batch_create" in names and "pq_run_sweeps" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_header_layout():
    # offsets of the ctypes mirror follow the C layout rules of the header
    assert C.sizeof(capi.ChainStats) == 8 * 20
    assert capi.ChainDesc.nsection.offset > capi.ChainDesc.sample.offset
    lib = capi.load_library()
    assert lib.pq_last_error() is not None
