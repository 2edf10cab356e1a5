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
    capi.check(_lib().pq_write_paths_file(fn.encode(), capi.dptr(beads), None, 2, 2, 2))
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
    perm = np.array([2, 0, 1, 4, 3], dtype=np.int32)
    capi.check(lib.pq_write_paths_file(fn.encode(), capi.dptr(beads), capi.iptr(perm), 16, 5, 3))
    assert open(fn).readline() == "#Permutations 2 0 1 4 3 \n"
    back = np.zeros_like(beads)
    got = np.zeros(5, dtype=np.int32)
    capi.check(lib.pq_read_paths_file(fn.encode(), capi.dptr(back), capi.iptr(got), 16, 5, 3, 1))
    assert np.allclose(back, beads, rtol=1e-14, atol=0)
    assert list(got) == list(perm)


def test_read_legacy_file_with_wrap_slice(tmp_path):
    """No permutation line: the slice after the last one closes the path and
    fixes the permutation (identity here); a header with model state text is
    skipped whole."""
    fn = str(tmp_path / "old")
    with open(fn, "w") as f:
        f.write("#Path coordinates: 2 particles in 1-d spin 0 1\n")
        f.write("1.5 -2 \n0.25 4 \n1.5 -2 \n")
    out = np.zeros(4)
    capi.check(_lib().pq_read_paths_file(fn.encode(), capi.dptr(out), None, 2, 2, 1, 1))
    assert list(out) == [1.5, -2.0, 0.25, 4.0]


def test_read_exchange(tmp_path):
    """An exchange is returned through perm and refused without it; a legacy
    file's wrap slice that swaps the particles is the same exchange."""
    fn = str(tmp_path / "perm")
    with open(fn, "w") as f:
        f.write("#Permutations 1 0 \n#Path coordinates: 2 particles in 1-d\n1 2 \n3 4 \n")
    out = np.zeros(4)
    perm = np.zeros(2, dtype=np.int32)
    assert _lib().pq_read_paths_file(fn.encode(), capi.dptr(out), None, 2, 2, 1, 1) == capi.PQ_ERR_ARG
    capi.check(_lib().pq_read_paths_file(fn.encode(), capi.dptr(out), capi.iptr(perm), 2, 2, 1, 1))
    assert list(perm) == [1, 0] and list(out) == [1, 2, 3, 4]
    with open(fn, "w") as f:
        f.write("#Path coordinates: 2 particles in 1-d\n1 2 \n3 4 \n2 1 \n")
    assert _lib().pq_read_paths_file(fn.encode(), capi.dptr(out), None, 2, 2, 1, 1) == capi.PQ_ERR_ARG
    perm[:] = 0
    capi.check(_lib().pq_read_paths_file(fn.encode(), capi.dptr(out), capi.iptr(perm), 2, 2, 1, 1))
    assert list(perm) == [1, 0]
    with open(fn, "w") as f:
        f.write("#Permutations 1 1 \n#Path coordinates: 2 particles in 1-d\n1 2 \n3 4 \n")
    assert _lib().pq_read_paths_file(fn.encode(), capi.dptr(out), capi.iptr(perm), 2, 2, 1, 1) == capi.PQ_ERR_ARG


def test_read_with_bfactor_relabels_copies(tmp_path):
    """ReadPaths puts each copy's last slice back with the file's permutation
    p, last copy first (PathReader.cc:113-117), so copy k carries p applied k
    times and the path permutation is p^bfactor.  A 3-cycle, bfactor 2: copy 1
    is relabelled once, bead(i) = file bead(p[i]), and the permutation is p^2."""
    fn = str(tmp_path / "cyc")
    with open(fn, "w") as f:
        f.write("#Permutations 1 2 0 \n#Path coordinates: 3 particles in 1-d\n10 11 12 \n20 21 22 \n")
    out = np.zeros((4, 3, 1))
    perm = np.zeros(3, dtype=np.int32)
    capi.check(_lib().pq_read_paths_file(fn.encode(), capi.dptr(out), capi.iptr(perm), 4, 3, 1, 2))
    assert out[:, :, 0].tolist() == [[10, 11, 12], [20, 21, 22], [11, 12, 10], [21, 22, 20]]
    assert list(perm) == [2, 0, 1]


def test_read_with_bfactor_tiles_the_file(tmp_path):
    """bfactor 2: the file's M/2 slices fill slices s and s + M/2."""
    fn = str(tmp_path / "b")
    with open(fn, "w") as f:
        f.write("#Permutations 0 \n#Path coordinates: 1 particles in 1-d\n7 \n8 \n")
    out = np.zeros(4)
    capi.check(_lib().pq_read_paths_file(fn.encode(), capi.dptr(out), None, 4, 1, 1, 2))
    assert list(out) == [7.0, 8.0, 7.0, 8.0]


def test_short_file_is_an_error(tmp_path):
    fn = str(tmp_path / "s")
    with open(fn, "w") as f:
        f.write("#Permutations 0 \n#Path coordinates: 1 particles in 1-d\n7 \n")
    out = np.zeros(4)
    assert _lib().pq_read_paths_file(fn.encode(), capi.dptr(out), None, 4, 1, 1, 1) == capi.PQ_ERR_ARG
