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
    assert C.sizeof(capi.ChainStats) == 8 * 23
    assert capi.ChainDesc.nsection.offset > capi.ChainDesc.sample.offset
    lib = capi.load_library()
    assert lib.pq_last_error() is not None


def test_ctypes_mirror_matches_the_c_layout(tmp_path):
    """sizeof and every field offset of the descriptor structs, as the C
    compiler lays out include/piqmc_hip.h, against the ctypes mirror."""
    import shutil
    import subprocess
    import pytest
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"pq_species_desc": capi.SpeciesDesc, "pq_sho_desc": capi.ShoDesc,
               "pq_coulomb_desc": capi.CoulombDesc, "pq_node_desc": capi.NodeDesc,
               "pq_sample_desc": capi.SampleDesc, "pq_section_group": capi.SectionGroup,
               "pq_chain_desc": capi.ChainDesc, "pq_chain_stats": capi.ChainStats,
               "pq_xml_input": capi.XmlInput}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "piqmc_hip.h"', 'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        cname, field, value = line.split()
        py = structs[cname]
        got = C.sizeof(py) if field == "size" else getattr(py, field).offset
        assert got == int(value), (cname, field, got, value)


def _create_rc(desc):
    lib = capi.load_library()
    h = C.c_void_p()
    arr = capi.descs_array([desc])
    return lib.pq_batch_create(arr, 1, capi.PQ_PREC_FP64, 0, C.byref(h)), h


def test_batch_create_rejects_empty_paths_and_bad_cells():
    """Descriptors are checked before any device call: a chain of zero
    slices, or a cell side that is zero, negative or not finite, is
    PQ_ERR_ARG on any host (no kernel runs)."""
    bad = []
    d = capi.make_desc(2, 8, 0, 0.1, [(1, 1.0, 0.0)], nsection=1)
    d.nslice = 0
    bad.append(d)
    for side in (0.0, -1.0, float("nan"), float("inf")):
        d = capi.make_desc(3, 16, 2, 0.1, [(2, 1.0, -1.0)], cell=4.0)
        d.cell[1] = side
        bad.append(d)
    d = capi.make_desc(2, 16, 2, 0.1, [(2, 1.0, -1.0)])
    d.tau = float("nan")
    bad.append(d)
    for d in bad:
        rc, h = _create_rc(d)
        assert rc == capi.PQ_ERR_ARG, (rc, capi.load_library().pq_last_error())
        assert not h.value
