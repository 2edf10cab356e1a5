"""The benchmark workloads at their real shapes, on the device against the
oracle: identical draws, accept/reject decisions, node errors and paths, and
bit-exact g(r) counts, under Philox streams and injected glibc streams.

* the 32 + 32 electron-hole box: FreeParticleNodes 32 x 32 per species
  (the 32-wide FP64 MFMA elimination), FreeMoverPBC with image draws, mass-2
  holes, minimum-image Coulomb tables on 1024 knots, L = 10, M = 1024, L = 5;
* eight chains of the 4096-chain 3D batch, N_s from 3 to 32 and M up to
  2048, in one launch and with the node-width split, with the 256-bin
  g(r) of the two species (SHONodes at a node temperature that FP64
  resolves; resolved_nodes says why);
* the 3+3-electron SHONodes dot at its full M = 400;
* eight of the heaviest chains of the 1024-chain 2D batch (M = 1024);
* one free particle on a periodic torus sampled by 256 device chains with
  FreeMoverPBC: the winding-number law and energy of the exact partition
  function."""
import math

import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi
from gpu_helpers import glibc_uniforms

pytestmark = pytest.mark.gpu

CFG5_PICK = (114, 361, 375, 633, 865, 92, 9, 1181)   # N 40, 47, 6, 10, 19, 64, 64, 57; M 2048 x 6


def resolved_nodes(descs, temperature=0.5):
    """The same chains with SHONodes at temperature 0.5 (omega beta_node / 2
    = 1 for omega = 1; the <FixedNodeAction temperature> attribute).  At the
    workloads' own node temperature 1/(tau M) the SHO density matrix of
    beta/2 up to 51 is rank one to working precision (its second singular
    value is ~1e-16 of the first already at N_s = 3; see
    test_oracle_pins.test_sho_node_matrix_rank_one_at_low_temperature), so
    every determinant is rounding noise and its sign -- hence the sampler's
    decisions -- differs between any two FP64 evaluations with different
    operation order, the reference's own included.  At temperature 0.5 the
    32 x 32 matrices have condition numbers below ~1e6 and the determinants
    keep ~10 digits, so identical decisions are a meaningful test; shapes,
    kernels and code paths are the workloads' own."""
    out = []
    for d in descs:
        e = capi.ChainDesc()
        capi.C.pointer(e)[0] = d
        for a in range(e.nnodes):
            if e.nodes[a].model == capi.PQ_NODES_SHO:
                e.nodes[a].temperature = temperature
        out.append(e)
    return out


def _run_pair(descs, nsweep, inject, seed=77, split=None, team=-1, precision=capi.PQ_PREC_FP64):
    b = capi.Batch(descs, precision=precision)
    if team != -1:
        b.set_schedule(team, -1)
    if split is not None:
        b.set_split(*split)
    ocs = []
    us = []
    for c, d in enumerate(descs):
        beads = capi.initial_beads(d, 3000 + c)
        b.set_paths(c, beads)
        oc = ol.Chain(d)
        oc.set_beads(beads)
        if inject:
            u = glibc_uniforms(seed + c, 600000)
            us.append(u)
            oc.rng_inject(u)
        else:
            oc.rng_philox(seed, c)
        ocs.append(oc)
    if inject:
        b.set_uniforms(np.stack(us))
    else:
        b.seed(seed, 0)
    b.run_sweeps(nsweep)
    b.sync()
    for oc in ocs:
        oc.run_sweeps(nsweep)
    return b, ocs


def _assert_same(b, ocs, tol=1e-9):
    for c, oc in enumerate(ocs):
        sg, so = b.stats(c), oc.stats()
        assert sg.draws == so.draws, (c, sg.draws, so.draws)
        assert list(sg.tries) == list(so.tries), c
        assert list(sg.accepts) == list(so.accepts), c
        assert sg.bead_updates == so.bead_updates, c
        assert sg.node_errors == so.node_errors, c
        assert np.max(np.abs(b.get_paths(c) - oc.get_beads())) < tol, c


@pytest.mark.parametrize("inject", [True, False])
def test_electron_hole_box(inject):
    """32 e (m 1) + 32 h (m 2), box L = 10, FreeMoverPBC, FreeParticleNodes
    32 x 32 on the 32-wide MFMA instance: two sweeps."""
    print(f"[test] electron-hole box inject={inject}", flush=True)
    d = capi.ehbox_desc()
    b, ocs = _run_pair([d], 2, inject, team=256)
    _assert_same(b, ocs)
    e_gpu, e_cpu = float(b.energy()[0]), ocs[0].energy()
    assert abs(e_gpu - e_cpu) <= 1e-12 * max(1.0, abs(e_cpu)), (e_gpu, e_cpu)
    st = b.stats(0)
    assert st.accepts[0] > 0


@pytest.mark.parametrize("split", [None, (-1, 256)])
def test_large_3d_batch_chains(split):
    """Eight chains of the 4096-chain batch (N_s 3..32, M up to 2048, SHONodes,
    256-bin g(r) to r = 8): two sweeps, in one launch or split at the node
    width (32-wide chains on their own launch)."""
    print(f"[test] 3D batch chains split={split}", flush=True)
    all_descs = capi.batch3d_descs(4096)
    descs = resolved_nodes([all_descs[c] for c in CFG5_PICK])
    assert max(d.nslice for d in descs) == 2048
    assert max(d.species[0].count for d in descs) == 32
    assert min(d.species[1].count for d in descs) == 3
    b, ocs = _run_pair(descs, 2, False, split=split)
    _assert_same(b, ocs)
    counts01 = b.pair_histogram(0, 1, 0.0, 8.0, 256)
    counts00 = b.pair_histogram(0, 0, 0.0, 8.0, 256)
    for c, oc in enumerate(ocs):
        assert np.array_equal(counts01[c], oc.pair_histogram(0, 1, 0.0, 8.0, 256)), c
        assert np.array_equal(counts00[c], oc.pair_histogram(0, 0, 0.0, 8.0, 256)), c
        assert counts01[c].sum() > 0
    e = b.energy()
    for c, oc in enumerate(ocs):
        assert abs(e[c] - oc.energy()) <= 1e-12 * max(1.0, abs(oc.energy())), c


def test_large_3d_batch_chains_injected():
    print("[test] 3D batch chains, injected glibc streams", flush=True)
    all_descs = capi.batch3d_descs(4096)
    descs = resolved_nodes([all_descs[c] for c in CFG5_PICK[:4]])
    b, ocs = _run_pair(descs, 1, True)
    _assert_same(b, ocs)


@pytest.mark.parametrize("inject", [True, False])
def test_three_plus_three_dot_full_size(inject):
    """The 3+3-electron SHONodes dot at M = 400, four sweeps."""
    print(f"[test] 3+3 dot M=400 inject={inject}", flush=True)
    d = capi.dot2d_3p3_desc()
    assert d.nslice == 400
    b, ocs = _run_pair([d], 4, inject)
    _assert_same(b, ocs)


def test_heaviest_2d_batch_chains():
    """Eight of the 1024-chain 2D batch's heaviest chains (M = 1024, N up to
    20, SHO and free-particle nodes), with the heavy chains split off."""
    print("[test] heaviest 2D batch chains", flush=True)
    all_descs = capi.batch2d_descs(1024)
    heavy = sorted(range(1024), key=lambda c: -all_descs[c].nslice * all_descs[c].npart ** 2)
    descs = resolved_nodes([all_descs[c] for c in heavy[:8]])
    assert all(d.nslice == 1024 for d in descs)
    b, ocs = _run_pair(descs, 2, False, split=(2, 512))
    _assert_same(b, ocs)


def test_free_mover_pbc_torus_on_device():
    """256 chains of one free particle on a 2D torus of side 1 (beta 0.2, M
    128, sections of the whole path): <w^2> per dimension and the thermal
    energy against Z = sum_w exp(-w^2 L^2/(4 lambda beta))."""
    print("[test] torus statistics on the device", flush=True)
    from test_oracle_pins import _torus_desc, torus_exact, windings
    w2_exact, e_exact = torus_exact()
    n = 256
    d = _torus_desc(7, 1)
    b = capi.Batch([d] * n)
    for c in range(n):
        b.set_paths(c, np.zeros((128, 1, 2)))
    b.seed(2026, 0)
    b.run_sweeps(500)
    ws, es = [], []
    for blk in range(40):
        e = b.run_measured(25, 1)
        es.append(e.mean(0))
        ws.append(np.array([windings(b.get_paths(c)) for c in range(n)]))
    ws = np.array(ws)
    w2 = float((ws * ws).mean())
    # chains are independent: the error of the mean over chains of their time averages
    per_chain = (ws * ws).mean(axis=(0, 2))
    err = per_chain.std(ddof=1) / math.sqrt(n)
    assert abs(w2 - w2_exact) < max(4 * err, 0.01), (w2, w2_exact, err)
    ec = np.array(es).mean(0)
    eerr = ec.std(ddof=1) / math.sqrt(n)
    assert abs(ec.mean() - e_exact) < 4 * eerr, (ec.mean(), eerr, e_exact)
