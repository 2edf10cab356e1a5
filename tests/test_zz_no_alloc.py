"""pq_action_difference_static, the entry point that allocates nothing,
against pq_action_difference through the C ABI: bitwise the same
differences on two calls in a row (the second with fewer queries, so stale
contents of the shared device array would show), and a clear refusal of
queries beyond its device array."""
import ctypes as C

import numpy as np
import pytest

import piqmc_capi as capi
from gpu_helpers import random_section

pytestmark = pytest.mark.gpu


def _queries(descs, nq, rng):
    chains, levels, secs, movs, idxs = [], [], [], [], []
    for q in range(nq):
        c = int(rng.integers(0, len(descs)))
        d = descs[c]
        sec, mov, idx = random_section(d, rng)
        chains.append(c); levels.append(q % (d.nlevel + 1)); idxs.append(idx)
        secs.append(sec.ravel()); movs.append(mov.ravel())
    return (np.array(chains, dtype=np.int32), np.array(levels, dtype=np.int32),
            np.ascontiguousarray(np.concatenate(secs)), np.ascontiguousarray(np.concatenate(movs)),
            np.array(idxs, dtype=np.int32))


def _call(fn, b, q):
    chains, levels, sec, mov, idx = q
    out = np.full(len(chains), np.nan)
    rc = fn(b.h, len(chains), capi.iptr(chains), capi.iptr(levels), capi.dptr(sec), capi.dptr(mov),
            capi.iptr(idx), 1, capi.dptr(out))
    return rc, out


def test_static_entry_point_matches_the_allocating_one():
    print("[test] no-alloc action difference", flush=True)
    lib = capi.load_library()
    stat = lib.pq_action_difference_static
    stat.restype = C.c_int
    stat.argtypes = lib.pq_action_difference.argtypes
    pbc = capi.make_desc(3, 64, 3, 0.05, [(2, 1.0, -1.0), (2, 2.0, 1.0)], cell=4.0,
                         coulomb=dict(epsilon=1.0, norder=3, rmin=0.0, rmax=0.0, ngpts=512),
                         mover=capi.PQ_MOVER_FREE_PBC)
    dot = capi.make_desc(3, 64, 4, 0.05, [(3, 1.0, -1.0), (2, 1.0, -1.0)], sho=[(0, 1.0), (1, 0.7)],
                         coulomb=dict(epsilon=1.0, norder=3, rmin=0.005, rmax=50.0, ngpts=1024))
    descs = [capi.dot3d_2e_desc(), pbc, dot]
    b = capi.Batch(descs)
    rng = np.random.default_rng(17)
    for nq in (64, 9):
        q = _queries(descs, nq, rng)
        rc, want = _call(lib.pq_action_difference, b, q)
        assert rc == capi.PQ_OK
        rc, got = _call(stat, b, q)
        assert rc == capi.PQ_OK, lib.pq_last_error()
        assert np.array_equal(got, want), nq
        assert np.all(np.isfinite(got))
    # 12000 queries of the first chain need ~7.9 MiB: refused, nothing launched
    chains = np.zeros(12000, dtype=np.int32)
    S, N, D = 9, 2, 3
    big = (chains, np.zeros(12000, dtype=np.int32), np.zeros(12000 * S * N * D), np.zeros(12000 * S * D),
           np.zeros(12000, dtype=np.int32))
    rc, _ = _call(stat, b, big)
    assert rc == capi.PQ_ERR_ARG
    assert b"split them" in lib.pq_last_error()
    # and the entry point still works after the refusal
    q = _queries(descs, 5, rng)
    assert np.array_equal(_call(stat, b, q)[1], _call(lib.pq_action_difference, b, q)[1])
