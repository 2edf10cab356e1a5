"""Path files of <WritePaths>/<ReadPaths> (src/algorithm/WritePaths.cc:36-112,
src/algorithm/PathReader.cc:17-132): host code of the C ABI, no device.
The expected text is the reference's stream output: "#Permutations " with
each entry followed by a space, the header line, then each slice on one line
with every coordinate followed by a space, at std::ostream precision 15
(printf %.15g)."""
import ctypes as C
import os

import numpy as np
import pytest

import piqmc_capi as capi

pytestmark = pytest.mark.skipif(not os.path.exists(capi.LIB_PATH), reason="library not built")


def _lib():
    return capi.load_library()


def test_write_matches_reference_layout(tmp_path):
    beads = np.array([[[0.5, -1.25], [3.0, 1e-05]],
                      [[1.0 / 3.0, 2.0e10], [-0.0625, 123456.789]]])
    fn = str(tmp_path / "paths")
    capi.check(_lib().pq_write_paths_file(fn.encode(), capi.dptr(beads), 2, 2, 2))
    want = ("#Permutations 0 1 \n"
            "#Path coordinates: 2 particles in 2-d\n"
            "0.5 -1.25 3 1e-05 \n"
            "0.333333333333333 20000000000 -0.0625 123456.789 \n")
    assert open(fn).read() == want


def test_round_trip_keeps_fifteen_digits(tmp_path):
    rng = np.random.default_rng(3)
    beads = rng.standard_normal((16, 5, 3)) * 10.0
    fn = str(tmp_path / "p")
    lib = _lib()
    capi.check(lib.pq_write_paths_file(fn.encode(), capi.dptr(beads), 16, 5, 3))
    back = np.zeros_like(beads)
    capi.check(lib.pq_read_paths_file(fn.encode(), capi.dptr(back), 16, 5, 3, 1))
    assert np.allclose(back, beads, rtol=1e-14, atol=0)


def test_read_legacy_file_with_wrap_slice(tmp_path):
    """No permutation line: the slice after the last one closes the path and
    fixes the permutation (identity here); a header with model state text is
    skipped whole."""
    fn = str(tmp_path / "old")
    with open(fn, "w") as f:
        f.write("#Path coordinates: 2 particles in 1-d spin 0 1\n")
        f.write("1.5 -2 \n0.25 4 \n1.5 -2 \n")
    out = np.zeros(4)
    capi.check(_lib().pq_read_paths_file(fn.encode(), capi.dptr(out), 2, 2, 1, 1))
    assert list(out) == [1.5, -2.0, 0.25, 4.0]


def test_read_refuses_exchange(tmp_path):
    fn = str(tmp_path / "perm")
    with open(fn, "w") as f:
        f.write("#Permutations 1 0 \n#Path coordinates: 2 particles in 1-d\n1 2 \n3 4 \n")
    out = np.zeros(4)
    assert _lib().pq_read_paths_file(fn.encode(), capi.dptr(out), 2, 2, 1, 1) == capi.PQ_ERR_ARG
    # a legacy file whose wrap slice swaps the particles is an exchange too
    with open(fn, "w") as f:
        f.write("#Path coordinates: 2 particles in 1-d\n1 2 \n3 4 \n2 1 \n")
    assert _lib().pq_read_paths_file(fn.encode(), capi.dptr(out), 2, 2, 1, 1) == capi.PQ_ERR_ARG


def test_read_with_bfactor_tiles_the_file(tmp_path):
    """bfactor 2: the file's M/2 slices fill slices s and s + M/2."""
    fn = str(tmp_path / "b")
    with open(fn, "w") as f:
        f.write("#Permutations 0 \n#Path coordinates: 1 particles in 1-d\n7 \n8 \n")
    out = np.zeros(4)
    capi.check(_lib().pq_read_paths_file(fn.encode(), capi.dptr(out), 4, 1, 1, 2))
    assert list(out) == [7.0, 8.0, 7.0, 8.0]


def test_short_file_is_an_error(tmp_path):
    fn = str(tmp_path / "s")
    with open(fn, "w") as f:
        f.write("#Permutations 0 \n#Path coordinates: 1 particles in 1-d\n7 \n")
    out = np.zeros(4)
    assert _lib().pq_read_paths_file(fn.encode(), capi.dptr(out), 4, 1, 1, 1) == capi.PQ_ERR_ARG
