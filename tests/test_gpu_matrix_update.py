"""FreeParticleNodes::MatrixUpdate (useUpdates) on the device against the oracle:

* the fixed-node protocol (initialize, per-level action difference, accept):
  action differences at 1e-11 and crossings exactly, with one or two moving
  particles of a species and one of each species, with and without the nodal
  action (acceptLastMove's two branches), on the row-per-lane and the 32-wide
  instances;
* sweeps: identical draws, decisions, node errors and paths on 2D and 3D
  boxes, with a two-particle Sample (the general sampler instances) and on the
  32 + 32 electron-hole box, under Philox and injected glibc streams."""
import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi
from gpu_helpers import rel_close
from test_gpu_configs import _assert_same, _run_pair

pytestmark = pytest.mark.gpu


def _box(D, counts, cell, nslice, nlevel, nodal=True, samples=None, masses=(1.0, 1.5)):
    sp = [(counts[k], masses[k], -1.0) for k in range(len(counts))]
    d = capi.make_desc(D, nslice, nlevel, 0.05, sp, cell=cell,
                       nodes=[(k, capi.PQ_NODES_FREE, 0.0, True) for k in range(len(counts))],
                       mover=capi.PQ_MOVER_FREE_PBC, samples=samples)
    for a in range(d.nnodes):
        d.nodes[a].nodal_action = 1 if nodal else 0
    return d


def _lattice_sections(d, rng, noise):
    """both sections near one configuration with every species on a simple
    lattice of the box, particle 1 near particle 0 (so the nodal action is
    not negligible)"""
    S = (1 << d.nlevel) + 1
    D, L = d.ndim, d.cell[0]
    base = np.zeros((1, d.npart, D))
    i = 0
    for s in range(d.nspecies):
        n = d.species[s].count
        side = int(np.ceil(n ** (1.0 / D)))
        for k in range(n):
            pos = [(k // side ** q) % side for q in range(D)]
            base[0, i] = (np.array(pos) + 0.5 + 0.3 * s) * (L / side) - L / 2
            i += 1
    base[0, 1] = base[0, 0] + 0.25 * (L / 4) * np.eye(D)[0]
    sec1 = base + noise * rng.standard_normal((S, d.npart, D))
    sec2 = base + noise * rng.standard_normal((S, d.npart, D))
    return np.ascontiguousarray(sec1), np.ascontiguousarray(sec2)


CASES = {
    "2d": dict(D=2, counts=(4, 3), cell=4.0, nslice=64, nlevel=3),
    "3d_12": dict(D=3, counts=(12, 8), cell=6.0, nslice=128, nlevel=4),
    "3d_20": dict(D=3, counts=(20, 5), cell=7.0, nslice=128, nlevel=4),
}


@pytest.mark.parametrize("nodal", [True, False])
@pytest.mark.parametrize("name", list(CASES))
def test_matrix_update_protocol(name, nodal):
    print(f"[test] matrix update protocol {name} nodal={nodal}", flush=True)
    d = _box(nodal=nodal, **CASES[name])
    b = capi.Batch([d])
    oc = ol.Chain(d)
    lib = capi.load_library()
    rng = np.random.default_rng(17)
    sec1, sec2 = _lattice_sections(d, rng, 0.03)
    capi.check(lib.pq_section_begin(b.h, 0, capi.dptr(sec1), capi.dptr(sec2)))
    oc.fn_initialize(sec1, sec2)
    n0 = d.species[0].count
    moves = [[1], [0, 2], [1, n0], [3], [1], [n0 + 1, 2]]
    naccept = 0
    for trial, idx in enumerate(moves * 2):
        mov = np.ascontiguousarray(sec1[:, idx, :].copy())
        mov[1:-1] += (0.02 if trial % 3 else 0.2) * rng.standard_normal(mov[1:-1].shape)
        ix = np.array(idx, dtype=np.int32)
        for level in range(d.nlevel, -1, -1):
            got = np.zeros(1)
            capi.check(lib.pq_double_action_difference(b.h, 0, level, capi.dptr(mov), capi.iptr(ix), len(idx),
                                                       capi.dptr(got)))
            want = oc.fn_action_difference(mov, idx, level)
            if want >= 1e100:
                assert got[0] == want, (trial, level)
                break
            assert rel_close(got[0], want, 1e-11, scale=1.0), (trial, level, got[0], want)
        else:
            capi.check(lib.pq_accept_last_move(b.h, 0))
            oc.fn_accept()
            sec1[:, idx, :] = mov
            naccept += 1
    assert naccept >= 4


SWEEP_CASES = {   # one batch per dimension
    2: [lambda: _box(2, (4, 3), 4.0, 64, 3),
        lambda: _box(2, (5, 3), 4.5, 64, 3, samples=[(0, 2), (1, 1)])],
    3: [lambda: _box(3, (12, 8), 6.0, 128, 4),
        lambda: _box(3, (20, 5), 7.0, 128, 4, nodal=False)],
}


@pytest.mark.parametrize("inject", [True, False])
@pytest.mark.parametrize("D", [2, 3])
def test_matrix_update_sweeps(D, inject):
    print(f"[test] matrix update sweeps D={D} inject={inject}", flush=True)
    descs = [f() for f in SWEEP_CASES[D]]
    b, ocs = _run_pair(descs, 3, inject)
    _assert_same(b, ocs)
    for c in range(len(descs)):
        assert b.stats(c).accepts[0] > 0, c


def test_matrix_update_electron_hole_box():
    """the 32 + 32 electron-hole box with useUpdates: two sweeps on the
    32-wide instance (initialize by MFMA elimination, moves by updates)"""
    print("[test] matrix update electron-hole box", flush=True)
    d = capi.ehbox_desc()
    for a in range(d.nnodes):
        d.nodes[a].use_updates = 1
    b, ocs = _run_pair([d], 2, False, team=256)
    _assert_same(b, ocs)
    assert b.stats(0).accepts[0] > 0
