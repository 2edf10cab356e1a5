"""GPU parity: the HIP library, called through its C ABI, against the oracle
on the same seeded inputs.  FP64 action and node values agree to 1e-12
relative (relative to the larger of the value and the summed term scale);
Metropolis decisions are identical given the same uniforms; integer
histograms are bit-exact."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi
from gpu_helpers import glibc_uniforms, random_section, rel_close, small_dot2d_3p3

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _descs():
    c1 = capi.dot3d_2e_desc()
    c2 = small_dot2d_3p3()
    pbc = capi.make_desc(3, 64, 3, 0.05, [(2, 1.0, -1.0), (2, 2.0, 1.0)], cell=4.0,
                         coulomb=dict(epsilon=1.0, norder=3, rmin=0.0, rmax=0.0, ngpts=512),
                         mover=capi.PQ_MOVER_FREE_PBC)
    # the reference's other choosers of one moving particle
    walk1 = capi.dot3d_2e_desc()
    for k in range(walk1.nsample):
        walk1.sample[k].chooser = capi.PQ_CHOOSER_WALKING
    all1 = capi.dot3d_2e_desc()
    all1.sample[0].chooser = capi.PQ_CHOOSER_ALL
    all1.sample[1].chooser = capi.PQ_CHOOSER_ALL_RANDOM
    walk2 = small_dot2d_3p3()
    for k in range(walk2.nsample):
        walk2.sample[k].chooser = capi.PQ_CHOOSER_WALKING
    return {"dot3d": c1, "dot2d": c2, "pbc": pbc, "dot3d_walk": walk1, "dot3d_all": all1, "dot2d_walk": walk2}


@pytest.mark.parametrize("name", ["dot3d", "dot2d", "pbc"])
def test_action_difference(name):
    print(f"[test] action_difference {name}", flush=True)
    desc = _descs()[name]
    b = capi.Batch([desc])
    oc = ol.Chain(desc)
    rng = np.random.default_rng(11)
    secs, movs, idxs, lev, want = [], [], [], [], []
    for q in range(48):
        sec, mov, idx = random_section(desc, rng)
        level = q % (desc.nlevel + 1)
        secs.append(sec); movs.append(mov); idxs.append(idx); lev.append(level)
        want.append(oc.action_difference(sec, mov, [idx], level))
    nq = len(secs)
    got = np.zeros(nq)
    lib = capi.load_library()
    sec_all = np.ascontiguousarray(np.concatenate([s.ravel() for s in secs]))
    mov_all = np.ascontiguousarray(np.concatenate([m.ravel() for m in movs]))
    ch = np.zeros(nq, dtype=np.int32)
    lv = np.array(lev, dtype=np.int32)
    ix = np.array(idxs, dtype=np.int32)
    capi.check(lib.pq_action_difference(b.h, nq, capi.iptr(ch), capi.iptr(lv), capi.dptr(sec_all),
                                        capi.dptr(mov_all), capi.iptr(ix), 1, capi.dptr(got)))
    want = np.array(want)
    assert rel_close(got, want, RTOL, scale=1.0), np.max(np.abs(got - want))


def test_pair_table_matches_oracle():
    import ctypes as C
    desc = _descs()["pbc"]
    b = capi.Batch([desc])
    oc = ol.Chain(desc)
    lib = capi.load_library()
    lib.pq_get_pair_table.restype = C.c_int
    for ip in range(ol.load().oq_npair(oc.h)):
        meta_o, grid_o, u_o, ut_o = oc.pair_table(ip)
        meta = np.zeros(4, dtype=np.int32)
        grid = np.zeros(2)
        u = np.zeros_like(u_o)
        ut = np.zeros_like(ut_o)
        capi.check(lib.pq_get_pair_table(b.h, 0, ip, capi.iptr(meta), capi.dptr(grid),
                                         capi.dptr(u), capi.dptr(ut)))
        assert list(meta) == list(meta_o)
        assert np.array_equal(grid, grid_o)
        assert np.allclose(u, u_o, rtol=1e-14, atol=0)
        assert np.allclose(ut, ut_o, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("model", [capi.PQ_NODES_SHO, capi.PQ_NODES_FREE])
def test_node_evaluate(model):
    print(f"[test] node_evaluate model={model}", flush=True)
    desc = capi.make_desc(2, 64, 4, 0.02, [(5, 1.0, -1.0), (3, 1.0, -1.0)], cell=6.0,
                          sho=[(0, 1.0), (1, 1.0)], nodes=[(0, model, 1.0), (1, model, 1.0)])
    b = capi.Batch([desc])
    oc = ol.Chain(desc)
    rng = np.random.default_rng(5)
    nmat = 9
    r1 = rng.uniform(-1.5, 1.5, size=(nmat, desc.npart, 2))
    r2 = r1 + 0.2 * rng.standard_normal(r1.shape)
    lib = capi.load_library()
    for inode, ns in ((0, 5), (1, 3)):
        det = np.zeros(nmat)
        err = np.zeros(nmat, dtype=np.int32)
        d1 = np.zeros(nmat * ns)
        d2 = np.zeros(nmat * ns)
        capi.check(lib.pq_node_evaluate(b.h, 0, inode, nmat, capi.dptr(np.ascontiguousarray(r1)),
                                        capi.dptr(np.ascontiguousarray(r2)), capi.dptr(det),
                                        capi.iptr(err), capi.dptr(d1), capi.dptr(d2)))
        for m in range(nmat):
            dw, ew = oc.node_evaluate(inode, r1[m], r2[m], 0)
            e1, e2 = oc.node_distance(inode, r1[m], r2[m], 0, ns)
            assert err[m] == ew
            assert rel_close(det[m], dw, RTOL, scale=0.0), (m, det[m], dw)
            assert rel_close(d1[m * ns:(m + 1) * ns], e1, 1e-11)
            assert rel_close(d2[m * ns:(m + 1) * ns], e2, 1e-11)


@pytest.mark.parametrize("ns,model,L", [(14, capi.PQ_NODES_SHO, 3), (24, capi.PQ_NODES_FREE, 3),
                                        (32, capi.PQ_NODES_SHO, 4)])
def test_large_species_nodes(ns, model, L):
    """Node matrices of 14, 24 and 32 particles (the 16- and 32-wide
    instances; at 32 particles and L = 4 the sampler's fixed-node state no
    longer fits the LDS budget and lives in global memory): determinants and
    distances against the oracle, then a few sampler sweeps with identical
    decisions."""
    print(f"[test] large_species_nodes ns={ns} model={model} L={L}", flush=True)
    desc = capi.make_desc(3, 64, L, 0.05, [(ns, 1.0, -1.0), (2, 1.0, -1.0)], cell=12.0,
                          sho=[(0, 1.0), (1, 1.0)], nodes=[(0, model, 1.0)], nsection=2)
    b = capi.Batch([desc])
    oc = ol.Chain(desc)
    rng = np.random.default_rng(ns)
    nmat = 5
    r1 = rng.uniform(-2.0, 2.0, size=(nmat, desc.npart, 3))
    r2 = r1 + 0.3 * rng.standard_normal(r1.shape)
    lib = capi.load_library()
    det = np.zeros(nmat)
    err = np.zeros(nmat, dtype=np.int32)
    d1 = np.zeros(nmat * ns)
    d2 = np.zeros(nmat * ns)
    capi.check(lib.pq_node_evaluate(b.h, 0, 0, nmat, capi.dptr(np.ascontiguousarray(r1)),
                                    capi.dptr(np.ascontiguousarray(r2)), capi.dptr(det),
                                    capi.iptr(err), capi.dptr(d1), capi.dptr(d2)))
    for m in range(nmat):
        dw, ew = oc.node_evaluate(0, r1[m], r2[m], 0)
        e1, e2 = oc.node_distance(0, r1[m], r2[m], 0, ns)
        assert err[m] == ew
        # larger matrices: the condition number scales the rounding of both LUs
        assert rel_close(det[m], dw, 1e-10, scale=0.0), (m, det[m], dw)
        assert rel_close(d1[m * ns:(m + 1) * ns], e1, 1e-9)
        assert rel_close(d2[m * ns:(m + 1) * ns], e2, 1e-9)
    bs, os_ = _paired(desc, 3, seed=11, team={32: 256, 24: 512}.get(ns, -1))
    sg, so = bs.stats(0), os_.stats()
    assert sg.draws == so.draws
    assert list(sg.accepts) == list(so.accepts)
    assert np.max(np.abs(bs.get_paths(0) - os_.get_beads())) < 1e-9


@pytest.mark.parametrize("many_body", [0, 1])
def test_fixed_node_protocol(many_body):
    """initialize, per-level action difference (with crossings) and accept on
    one chain; the nodal action per particle or, with useManyBodyDistance,
    per slice from the many-body distance."""
    print(f"[test] fixed_node_protocol many_body={many_body}", flush=True)
    desc = small_dot2d_3p3(64)
    for a in range(desc.nnodes):
        desc.nodes[a].many_body_distance = many_body
    b = capi.Batch([desc])
    oc = ol.Chain(desc)
    lib = capi.load_library()
    rng = np.random.default_rng(3)
    S = (1 << desc.nlevel) + 1
    base = rng.standard_normal((1, desc.npart, 2))
    sec1 = np.ascontiguousarray(base + 0.15 * rng.standard_normal((S, desc.npart, 2)))
    sec2 = np.ascontiguousarray(base + 0.15 * rng.standard_normal((S, desc.npart, 2)))
    capi.check(lib.pq_section_begin(b.h, 0, capi.dptr(sec1), capi.dptr(sec2)))
    oc.fn_initialize(sec1, sec2)
    nbig = 0
    for trial in range(16):
        idx = int(rng.integers(0, desc.npart))
        mov = np.ascontiguousarray(sec1[:, idx:idx + 1, :].copy())
        if trial % 4 == 3:
            # walk onto a same-species partner: the determinant passes through zero
            other = (idx // 3) * 3 + (idx + 1) % 3
            mov[1:-1, 0] = sec1[1:-1, other] + 1e-3 * rng.standard_normal((S - 2, 2))
        else:
            mov[1:-1] += (0.05 if trial % 4 else 0.6) * rng.standard_normal((S - 2, 1, 2))
        for level in range(desc.nlevel, -1, -1):
            got = np.zeros(1)
            ix = np.array([idx], dtype=np.int32)
            capi.check(lib.pq_double_action_difference(b.h, 0, level, capi.dptr(mov), capi.iptr(ix), 1,
                                                       capi.dptr(got)))
            want = oc.fn_action_difference(mov, [idx], level)
            if want >= 1e100:
                nbig += 1
                assert got[0] == want
                break
            assert rel_close(got[0], want, 1e-11, scale=1.0), (trial, level, got[0], want)
        else:
            capi.check(lib.pq_accept_last_move(b.h, 0))
            oc.fn_accept()
            sec1[:, idx:idx + 1, :] = mov
    assert nbig >= 1


def _paired(desc, nsweep, seed=7, precision=capi.PQ_PREC_FP64, inject=False, team=-1, inv=-1,
            staging=(-1, -1)):
    b = capi.Batch([desc], precision=precision)
    b.set_schedule(team, inv)
    b.set_staging(*staging)
    oc = ol.Chain(desc)
    beads = capi.initial_beads(desc, seed)
    b.set_paths(0, beads)
    oc.set_beads(beads)
    if inject:
        u = glibc_uniforms(seed, 400000)
        b.set_uniforms(u[None, :])
        oc.rng_inject(u)
    else:
        b.seed(1234, 0)
        oc.rng_philox(1234, 0)
    b.run_sweeps(nsweep)
    b.sync()
    oc.run_sweeps(nsweep)
    return b, oc


@pytest.mark.parametrize("name,inject,team,inv", [("dot3d", True, -1, -1), ("dot3d", False, 128, -1),
                                                   ("dot2d", True, 64, 0), ("dot2d", True, 128, 1),
                                                   ("dot2d", False, 64, 1), ("pbc", False, -1, -1),
                                                   ("dot3d_walk", True, -1, -1), ("dot3d_all", True, -1, -1),
                                                   ("dot2d_walk", False, -1, -1), ("dot2d", True, 256, 0),
                                                   ("pbc", True, 256, -1), ("dot2d", False, 512, 1)])
def test_sampler_decisions(name, inject, team, inv):
    """Same uniforms in, same accept/reject decisions and paths out, for every schedule."""
    print(f"[test] sampler_decisions {name} inject={inject} team={team} inv={inv}", flush=True)
    desc = _descs()[name]
    b, oc = _paired(desc, 6, inject=inject, team=team, inv=inv)
    sg, so = b.stats(0), oc.stats()
    assert sg.draws == so.draws
    assert list(sg.tries) == list(so.tries)
    assert list(sg.accepts) == list(so.accepts)
    assert sg.bead_updates == so.bead_updates
    assert sg.node_errors == so.node_errors
    pg, po = b.get_paths(0), oc.get_beads()
    assert np.max(np.abs(pg - po)) < 1e-9


@pytest.mark.parametrize("name", ["dot3d", "dot2d"])
@pytest.mark.parametrize("staging", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_staging_keeps_decisions(name, staging):
    """Pair tables and node state in LDS or global memory: same decisions."""
    print(f"[test] staging {name} {staging}", flush=True)
    desc = _descs()[name]
    b, oc = _paired(desc, 4, inject=True, staging=staging)
    got = b.staging()
    assert got[0] == (staging[0] == 1) or desc.coulomb.enabled == 0
    sg, so = b.stats(0), oc.stats()
    assert sg.draws == so.draws
    assert list(sg.accepts) == list(so.accepts)
    assert np.max(np.abs(b.get_paths(0) - oc.get_beads())) < 1e-9


def test_autotune_keeps_decisions():
    """The autotuner's timing sweeps are ordinary sweeps: after it, the chain
    equals the oracle's after the same number of sweeps."""
    print("[test] autotune", flush=True)
    desc = small_dot2d_3p3(64)
    b = capi.Batch([desc])
    oc = ol.Chain(desc)
    beads = capi.initial_beads(desc, 8)
    b.set_paths(0, beads)
    oc.set_beads(beads)
    b.seed(4242, 0)
    oc.rng_philox(4242, 0)
    ms, nsw = b.autotune(1)
    assert ms > 0 and nsw >= 2
    b.run_sweeps(2)
    b.sync()
    draws = b.stats(0).draws
    for _ in range(64):
        if oc.stats().draws >= draws:
            break
        oc.run_sweeps(1)
    sg, so = b.stats(0), oc.stats()
    assert sg.draws == so.draws
    assert list(sg.accepts) == list(so.accepts)
    assert np.max(np.abs(b.get_paths(0) - oc.get_beads())) < 1e-9


def test_energy_estimator():
    print("[test] energy_estimator", flush=True)
    descs = [capi.dot3d_2e_desc(), capi.dot3d_2e_desc()]
    b = capi.Batch(descs)
    for c, d in enumerate(descs):
        beads = capi.initial_beads(d, 100 + c) * 0.7
        b.set_paths(c, beads)
        oc = ol.Chain(d)
        oc.set_beads(beads)
        want = oc.energy()
        got = b.energy()[c]
        assert rel_close(got, want, RTOL, scale=10.0), (got, want)


def test_histogram_bit_exact():
    print("[test] histogram", flush=True)
    desc = small_dot2d_3p3(128)
    b = capi.Batch([desc, desc])
    lib = capi.load_library()
    import ctypes as C
    for s1, s2 in ((0, 1), (0, 0)):
        counts = np.zeros(2 * 64, dtype=np.uint64)
        for c in range(2):
            beads = capi.initial_beads(desc, 40 + c, replicate=False) * 1.3
            b.set_paths(c, beads)
        capi.check(lib.pq_pair_histogram(b.h, s1, s2, 0.0, 4.0, 64,
                                         counts.ctypes.data_as(C.POINTER(C.c_uint64))))
        for c in range(2):
            oc = ol.Chain(desc)
            oc.set_beads(b.get_paths(c))
            want = oc.pair_histogram(s1, s2, 0.0, 4.0, 64)
            assert np.array_equal(counts[c * 64:(c + 1) * 64], want)


@pytest.mark.parametrize("split", [(0, 256), (3, 256), (2, 512)])
def test_batched_uneven_decisions(split):
    """A slice of the uneven batch (SHO and free-particle nodes, N 2..20, M up
    to 256), whole or with its heaviest chains in a launch of their own."""
    print(f"[test] batched_uneven split={split}", flush=True)
    descs = [d for d in capi.batch2d_descs(24) if d.nslice <= 256][:8]
    b = capi.Batch(descs)
    assert b.schedule()[0] == 128
    b.set_split(*split)
    assert b.split() == split
    ocs = []
    for c, d in enumerate(descs):
        beads = capi.initial_beads(d, 500 + c)
        b.set_paths(c, beads)
        oc = ol.Chain(d)
        oc.set_beads(beads)
        oc.rng_philox(99, c)
        ocs.append(oc)
    b.seed(99, 0)
    b.run_sweeps(2)
    b.sync()
    for c, oc in enumerate(ocs):
        oc.run_sweeps(2)
        sg, so = b.stats(c), oc.stats()
        assert sg.draws == so.draws, c
        assert list(sg.accepts) == list(so.accepts), c
        assert np.max(np.abs(b.get_paths(c) - oc.get_beads())) < 1e-9


def test_queued_heavy_launch_takes_every_chain():
    """The heavy chains' launch with 512-thread teams takes its chains from a
    queue with at most the resident workgroups (one per CU): 300 heavy chains
    are more than there are CUs, so workgroups run several chains in turn.
    Every chain, queued or not, makes the oracle's decisions."""
    print("[test] queued heavy launch", flush=True)
    rng = np.random.default_rng(5)
    descs = []
    for c in range(340):
        nup, ndn = int(rng.integers(1, 4)), int(rng.integers(1, 4))
        descs.append(capi.make_desc(2, 16 * int(rng.integers(1, 3)), 2, 0.05, [(nup, 1.0, -1.0), (ndn, 1.0, -1.0)],
                                    sho=[(0, 1.0), (1, 1.0)],
                                    nodes=[(0, capi.PQ_NODES_SHO, 1.0), (1, capi.PQ_NODES_SHO, 1.0)]))
    b = capi.Batch(descs)
    b.set_split(300, 512)
    assert b.split() == (300, 512)
    ocs = []
    for c, d in enumerate(descs):
        beads = capi.initial_beads(d, 900 + c, replicate=True)
        b.set_paths(c, beads)
        oc = ol.Chain(d)
        oc.set_beads(beads)
        oc.rng_philox(31, c)
        ocs.append(oc)
    b.seed(31, 0)
    b.run_sweeps(2)
    b.sync()
    paths = np.empty(int(sum(d.nbeads() for d in descs)))
    capi.check(b.lib.pq_get_all_paths(b.h, capi.dptr(paths), paths.size))
    off = 0
    for c, oc in enumerate(ocs):
        oc.run_sweeps(2)
        sg, so = b.stats(c), oc.stats()
        assert sg.sweeps == 2, c
        assert sg.draws == so.draws, c
        assert list(sg.accepts) == list(so.accepts), c
        n = descs[c].nbeads()
        assert np.max(np.abs(paths[off:off + n] - oc.get_beads().ravel())) < 1e-9, c
        off += n


def test_walking_chooser_veto_on_device():
    """An injected uniform of exactly 1.0 in WalkingChooser's veto slot turns
    every move down on the device as in the oracle; just below 1 the move runs."""
    print("[test] walking_veto", flush=True)
    d = capi.make_desc(2, 8, 2, 0.1, [(3, 1.0, 0.0), (2, 1.0, 0.0)],
                       samples=[(1, 1, capi.PQ_CHOOSER_WALKING)], nsection=1)
    beads = capi.initial_beads(d, 3)
    for veto, want_draws in ((1.0, 3), (0.999999, 12)):
        u = np.array([0.3, 0.75, veto] + [1e-6] * 64)
        b = capi.Batch([d])
        b.set_paths(0, beads)
        b.set_uniforms(u[None, :])
        b.run_sweeps(1)
        b.sync()
        oc = ol.Chain(d)
        oc.set_beads(beads)
        oc.rng_inject(u)
        oc.run_sweeps(1)
        sg, so = b.stats(0), oc.stats()
        assert sg.draws == so.draws == want_draws
        assert list(sg.tries) == list(so.tries)
        assert np.max(np.abs(b.get_paths(0) - oc.get_beads())) < 1e-12


def test_node_records_follow_new_paths():
    """The sampler reuses node determinants and distances of slice pairs
    across sections; replacing the paths must drop them.  Sweeps, new paths,
    more sweeps: decisions and paths stay the oracle's."""
    print("[test] node_records_follow_new_paths", flush=True)
    desc = small_dot2d_3p3(64)
    b = capi.Batch([desc])
    oc = ol.Chain(desc)
    for k, seed in enumerate((21, 22)):
        beads = capi.initial_beads(desc, seed)
        b.set_paths(0, beads)
        oc.set_beads(beads)
        if k == 0:
            b.seed(555, 0)
            oc.rng_philox(555, 0)
        b.run_sweeps(3)
        b.sync()
        oc.run_sweeps(3)
        sg, so = b.stats(0), oc.stats()
        assert sg.draws == so.draws
        assert list(sg.accepts) == list(so.accepts)
        assert np.max(np.abs(b.get_paths(0) - oc.get_beads())) < 1e-9


def test_mixed_precision_energy():
    """Mixed precision on the same Philox streams as FP64.  Both precisions
    move beads in FP64 and estimate in FP64, so a chain whose decisions never
    differed ends on the same path with the same energy, bit for bit; flips
    need a uniform within ~1e-6 of the acceptance probability, so nearly
    every chain keeps the FP64 decision stream over a few sweeps.  (Whether
    the flips bias the sampled energy is the statistical tests' question:
    test_mixed_precision_sampled_energy and the mixed cases of the SHO and
    fermion energies against their exact values.)"""
    print("[test] mixed_precision_energy", flush=True)
    descs = capi.batch2d_descs(64)
    runs = []
    for prec in (capi.PQ_PREC_FP64, capi.PQ_PREC_MIXED):
        b = capi.Batch(descs, precision=prec)
        for c, d in enumerate(descs):
            b.set_paths(c, capi.initial_beads(d, 1000 + c))
        b.seed(77, 0)
        b.run_sweeps(4)
        b.sync()
        runs.append((b.energy(), [b.get_paths(c) for c in range(len(descs))]))
    (e64, p64), (e32, p32) = runs
    same = np.array([np.array_equal(a, c) for a, c in zip(p64, p32)])
    print("identical chains", int(same.sum()), "of", len(descs))
    assert np.all(np.isfinite(e32))
    assert same.sum() >= 0.9 * len(descs)
    # identical decision streams give identical energies
    assert np.array_equal(e32[same], e64[same])


def test_mixed_precision_sampled_energy():
    """The interacting 2-electron 3D dot (Coulomb pair action, M 200, L 3)
    sampled on 512 independent chains in each precision, on different
    Philox streams: the two ensembles' mean thermal energies agree within 3
    combined standard errors -- no bias of the mixed mode's FP32 spring, SHO
    and pair-spline arithmetic at the ~0.5 % (0.019 Ha*) level of these runs."""
    print("[test] mixed_precision_sampled_energy", flush=True)
    d = capi.dot3d_2e_desc()
    runs = {}
    for prec, seed in ((capi.PQ_PREC_FP64, 61), (capi.PQ_PREC_MIXED, 62)):
        m = _chain_means(d, 512, 200, 200, seed, lambda c: capi.initial_beads(d, 9000 + c), precision=prec)
        runs[prec] = (m.mean(), m.std(ddof=1) / np.sqrt(len(m)))
    (e64, s64), (e32, s32) = runs[capi.PQ_PREC_FP64], runs[capi.PQ_PREC_MIXED]
    z = (e32 - e64) / np.hypot(s64, s32)
    print("E fp64", e64, "+-", s64, "mixed", e32, "+-", s32, "z", z)
    assert abs(z) < 3


@pytest.mark.parametrize("name", ["dot3d", "dot2d", "pbc"])
def test_action_difference_mixed(name):
    """FP32 link arithmetic with compensated sums against the FP64 oracle.
    Each of the ~10^2 summed terms carries a 2^-24 relative rounding, and
    dS cancels new against old, so the bound is 5e-5 relative to the larger
    of |dS| and 1 (FP64 meets 1e-12 on the same queries)."""
    print(f"[test] action_difference_mixed {name}", flush=True)
    desc = _descs()[name]
    b = capi.Batch([desc], precision=capi.PQ_PREC_MIXED)
    oc = ol.Chain(desc)
    rng = np.random.default_rng(12)
    secs, movs, idxs, lev, want = [], [], [], [], []
    for q in range(48):
        sec, mov, idx = random_section(desc, rng)
        level = q % (desc.nlevel + 1)
        secs.append(sec); movs.append(mov); idxs.append(idx); lev.append(level)
        want.append(oc.action_difference(sec, mov, [idx], level))
    nq = len(secs)
    got = np.zeros(nq)
    lib = capi.load_library()
    capi.check(lib.pq_action_difference(
        b.h, nq, capi.iptr(np.zeros(nq, dtype=np.int32)), capi.iptr(np.array(lev, dtype=np.int32)),
        capi.dptr(np.ascontiguousarray(np.concatenate([x.ravel() for x in secs]))),
        capi.dptr(np.ascontiguousarray(np.concatenate([m.ravel() for m in movs]))),
        capi.iptr(np.array(idxs, dtype=np.int32)), 1, capi.dptr(got)))
    want = np.array(want)
    err = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
    print("max relative error", float(err.max()))
    assert err.max() <= 5e-5


def _chain_means(desc, nchain, nequil, nmeasure, seed, beads_fn, virial_window=0, precision=capi.PQ_PREC_FP64):
    """Device sampler on nchain independent copies (Philox stream per chain):
    nequil sweeps, then nmeasure sweeps with one energy measurement each;
    returns the per-chain time averages (and of the virial estimator when a
    window is given)."""
    b = capi.Batch([desc] * nchain, precision=precision)
    for c in range(nchain):
        b.set_paths(c, beads_fn(c))
    b.seed(seed, 0)
    b.run_sweeps(nequil)
    acc = np.zeros(nchain)
    vir = np.zeros(nchain)
    for _ in range(nmeasure):
        b.run_sweeps(1)
        acc += b.energy()
        if virial_window:
            vir += b.virial_energy(virial_window)
    if virial_window:
        return acc / nmeasure, vir / nmeasure
    return acc / nmeasure


@pytest.mark.parametrize("precision", [capi.PQ_PREC_FP64, capi.PQ_PREC_MIXED])
def test_sampled_sho_energy_matches_analytic(precision):
    """Two distinguishable particles in a 3D dot (omega 1, beta 10, M 200,
    L 3, no Coulomb): E = 2 (3/2) coth(beta/2) = 3 coth(5).  Chains start on
    straight paths (one Gaussian position per particle) and equilibrate for
    300 sweeps, long against the relaxation of the path's slow modes; chain
    means of 256 independent chains agree with the exact value within 3
    standard errors, in FP64 and with the mixed mode's FP32 spring and SHO
    arithmetic in the sampler."""
    print(f"[test] sampled_sho_energy precision={precision}", flush=True)
    d = capi.make_desc(3, 200, 3, 0.05, [(1, 1.0, -1.0), (1, 1.0, -1.0)], sho=[(0, 1.0), (1, 1.0)])
    means, vmeans = _chain_means(d, 256, 300, 300, 31, lambda c: capi.initial_beads(d, 7000 + c, replicate=True),
                                 virial_window=10, precision=precision)
    exact = 3.0 / np.tanh(5.0)
    for name, m in (("thermo", means), ("virial", vmeans)):
        err = m.std(ddof=1) / np.sqrt(len(m))
        z = (m.mean() - exact) / err
        print(name, "E", m.mean(), "+-", err, "exact", exact, "z", z)
        assert abs(z) < 3
    # the windowed virial estimator has the smaller variance
    assert vmeans.std() < means.std()


@pytest.mark.parametrize("precision", [capi.PQ_PREC_FP64, capi.PQ_PREC_MIXED])
def test_sampled_fermion_energy_matches_exact_nodes(precision):
    """Two spin-polarised free fermions in a 2D dot with exact SHO nodes
    (beta 2, M 64, L 3): E = -d ln Z_F / d beta, Z_F = (Z1(b)^2 - Z1(2b))/2,
    Z1 = (2 sinh(b/2))^-2.  300 sweeps of equilibration from straight paths,
    then chain means of 256 chains within 3 standard errors, in both
    precision modes."""
    import math
    print(f"[test] sampled_fermion_energy precision={precision}", flush=True)
    beta, M = 2.0, 64
    d = capi.make_desc(2, M, 3, beta / M, [(2, 1.0, -1.0)], sho=[(0, 1.0)],
                       nodes=[(0, capi.PQ_NODES_SHO, 1.0)], nsection=M // 8)
    start = np.broadcast_to(np.array([[0.7, 0.0], [-0.7, 0.1]]), (M, 2, 2)).copy()
    means = _chain_means(d, 256, 300, 300, 47, lambda c: start, precision=precision)

    def lnz(bb):
        z1 = lambda x: (2 * math.sinh(x / 2)) ** -2
        return math.log(0.5 * (z1(bb) ** 2 - z1(2 * bb)))
    h = 1e-5
    exact = -(lnz(beta + h) - lnz(beta - h)) / (2 * h)
    err = means.std(ddof=1) / np.sqrt(len(means))
    z = (means.mean() - exact) / err
    print("E", means.mean(), "+-", err, "exact", exact, "z", z)
    assert abs(z) < 3


def test_batch_path_files_round_trip(tmp_path):
    """pq_write_all_paths / pq_read_all_paths: one file per clone (name + clone
    index when the batch has several), sampled paths back to 15 digits."""
    print("[test] batch_path_files", flush=True)
    d = capi.dot3d_2e_desc()
    b = capi.Batch([d, d])
    for c in range(2):
        b.set_paths(c, capi.initial_beads(d, 60 + c))
    b.seed(5, 0)
    b.run_sweeps(2)
    b.sync()
    before = [b.get_paths(c) for c in range(2)]
    base = str(tmp_path / "paths.out")
    lib = capi.load_library()
    capi.check(lib.pq_write_all_paths(b.h, base.encode()))
    assert os.path.exists(base + "0") and os.path.exists(base + "1")
    b2 = capi.Batch([d, d])
    capi.check(lib.pq_read_all_paths(b2.h, base.encode(), 1))
    for c in range(2):
        assert np.allclose(b2.get_paths(c), before[c], rtol=1e-14, atol=1e-300)


def test_measured_graph_matches_eager():
    """pq_run_measured's captured graph (heaviest chains forked onto the
    second stream), replayed several times with eager sweeps in between,
    gives the energies (and paths) of the same sweeps run one launch at a
    time on a twin batch."""
    print("[test] measured_graph", flush=True)
    descs = [d for d in capi.batch2d_descs(16) if d.nslice <= 256][:6]
    twins = []
    for _ in range(2):
        b = capi.Batch(descs, precision=capi.PQ_PREC_MIXED)
        for c, d in enumerate(descs):
            b.set_paths(c, capi.initial_beads(d, 300 + c))
        b.seed(17, 0)
        twins.append(b)
    g, e = twins
    g.set_split(2, 256)
    lib = capi.load_library()
    nblock, nsweep = 3, 2
    for rep in range(3):
        out = np.zeros(nblock * len(descs))
        capi.check(lib.pq_run_measured(g.h, nblock, nsweep, capi.dptr(out)))
        want = []
        for _ in range(nblock):
            e.run_sweeps(nsweep)
            want.append(e.energy())
        assert np.array_equal(out, np.concatenate(want)), rep
        # eager work between replays
        g.run_sweeps(1)
        e.run_sweeps(1)
    for c in range(len(descs)):
        assert np.array_equal(g.get_paths(c), e.get_paths(c))


def test_coulomb_energy_estimator():
    """CoulombEnergyEstimator on the device against the oracle (1e-12), open
    and periodic cells, after sweeps."""
    print("[test] coulomb energy estimator", flush=True)
    descs = [_descs()["dot3d"], _descs()["pbc"], small_dot2d_3p3()]
    for desc in descs:
        b, oc = _paired(desc, 3)
        for eps in (1.0, 12.5):
            got = b.coulomb_energy(eps)[0]
            assert rel_close(got, oc.coulomb_energy(eps), 1e-12, scale=1e-300), (got, oc.coulomb_energy(eps))


def test_virial_energy_estimator():
    """VirialEnergyEstimator on the device against the oracle for several
    windows, open and periodic cells, node chains (zero node force) and
    permuted bosons (1e-11 of the summed scale)."""
    print("[test] virial energy estimator", flush=True)
    bos = capi.make_desc(2, 64, 3, 2.0 / 64, [(4, 1.0, 0.0)], sho=[(0, 1.0)], cell=20.0,
                         samples=[(0, 2, capi.PQ_CHOOSER_WALKING), (0, 1, capi.PQ_CHOOSER_WALKING)])
    for desc in (_descs()["dot3d"], _descs()["pbc"], small_dot2d_3p3(), bos):
        b, oc = _paired(desc, 4)
        for nw in (1, 3, 8):
            got = b.virial_energy(nw)[0]
            want = oc.virial_energy(nw)
            assert rel_close(got, want, 1e-11, scale=1.0), (nw, got, want)
        assert rel_close(b.virial_energy(1)[0], b.energy()[0], 1e-11, scale=1.0)


def test_many_body_distance_sampler_decisions():
    """useManyBodyDistance in the device sampler: the oracle's decisions and
    paths for the 3+3 fermion dot."""
    print("[test] many-body distance sampler", flush=True)
    desc = small_dot2d_3p3(64)
    for a in range(desc.nnodes):
        desc.nodes[a].many_body_distance = 1
    b, oc = _paired(desc, 6)
    sg, so = b.stats(0), oc.stats()
    assert sg.draws == so.draws and list(sg.accepts) == list(so.accepts)
    assert np.max(np.abs(b.get_paths(0) - oc.get_beads())) < 1e-9


@pytest.mark.parametrize("team,heavy_team", [(128, 256), (64, 512)])
def test_node_width_split_decisions(team, heavy_team):
    """pq_set_split(-1): chains with 17-32-particle node species on the
    32-wide instance, the narrower ones on their own instance and layout;
    every chain makes the oracle's decisions."""
    print(f"[test] node-width split team={team} heavy_team={heavy_team}", flush=True)
    def chain(nup, ndn, M=64):
        return capi.make_desc(2, M, 3, 0.05, [(nup, 1.0, -1.0), (ndn, 1.0, -1.0)], sho=[(0, 1.0), (1, 1.0)],
                              nodes=[(0, capi.PQ_NODES_SHO, 1.0), (1, capi.PQ_NODES_SHO, 1.0)], nsection=8)
    descs = [chain(3, 2), chain(18, 3), chain(2, 2), chain(4, 20), chain(5, 5)]
    b = capi.Batch(descs)
    b.set_schedule(team, -1)
    capi.check(b.lib.pq_set_split(b.h, -1, heavy_team))
    h, ht = C.c_int(), C.c_int()
    capi.check(b.lib.pq_get_split(b.h, C.byref(h), C.byref(ht)))
    assert h.value == -1 and ht.value == heavy_team
    ocs = []
    for c, d in enumerate(descs):
        beads = capi.initial_beads(d, 40 + c, replicate=True)
        b.set_paths(c, beads)
        oc = ol.Chain(d)
        oc.set_beads(beads)
        oc.rng_philox(77, c)
        ocs.append(oc)
    b.seed(77, 0)
    b.run_sweeps(4)
    b.sync()
    for c, oc in enumerate(ocs):
        oc.run_sweeps(4)
        sg, so = b.stats(c), oc.stats()
        assert sg.draws == so.draws and list(sg.accepts) == list(so.accepts), c
        assert np.max(np.abs(b.get_paths(c) - oc.get_beads())) < 1e-9, c


def _dr_descs():
    """<Sample delayedRejection defaultSigmaFactor newSigmaFactor> cases."""
    wide = {"delayed_rejection": 1, "sigma_factor": 1.5, "new_sigma_factor": 3.0}
    sho1 = capi.make_desc(3, 64, 3, 1 / 64, [(1, 1.0, -1.0)], cell=50.0, sho=[(0, 1.0)],
                          samples=[(0, 1, capi.PQ_CHOOSER_SPECIES, 1, wide)])
    dot3d = capi.dot3d_2e_desc()
    for k in range(dot3d.nsample):
        dot3d.sample[k].delayed_rejection = 1
        dot3d.sample[k].sigma_factor = 1.2
    dot2d = small_dot2d_3p3()
    for k in range(dot2d.nsample):
        dot2d.sample[k].delayed_rejection = 1
        dot2d.sample[k].sigma_factor = 1.3
        dot2d.sample[k].new_sigma_factor = 0.6
    # a small box: the movers' periodic Gaussians at the delayed level
    box = capi.make_desc(3, 64, 3, 0.05, [(2, 1.0, -1.0), (2, 2.0, 1.0)], cell=2.0,
                         coulomb=dict(epsilon=1.0, norder=3, rmin=0.0, rmax=0.0, ngpts=512),
                         samples=[(0, 1, capi.PQ_CHOOSER_SPECIES, 1, wide), (1, 1, capi.PQ_CHOOSER_SPECIES, 1, wide)])
    # two moving particles per move (the multi-particle sampler instance)
    multi = small_dot2d_3p3()
    multi.nsample = 1
    multi.sample[0].nmoving = 2
    multi.sample[0].delayed_rejection = 1
    multi.sample[0].sigma_factor = 1.3
    # a scaled mover without delayed rejection
    fac = capi.dot3d_2e_desc()
    for k in range(fac.nsample):
        fac.sample[k].sigma_factor = 0.8
    return {"sho1": sho1, "dot3d": dot3d, "dot2d": dot2d, "box": box, "multi": multi, "factor": fac}


@pytest.mark.parametrize("name,inject", [("sho1", True), ("sho1", False), ("dot3d", True), ("dot2d", False),
                                         ("box", False), ("multi", True), ("factor", False)])
def test_delayed_rejection_decisions(name, inject):
    """Delayed rejection at level L-1 and the movers' sigma factors
    (MultiLevelSampler.cc:100-177, DoubleMLSampler.cc:122-245): the second
    proposal's draws, the lower levels proposed again after it, and the
    oracle's decisions, counters and paths."""
    print(f"[test] delayed rejection {name} inject={inject}", flush=True)
    desc = _dr_descs()[name]
    b, oc = _paired(desc, 6, inject=inject)
    sg, so = b.stats(0), oc.stats()
    assert sg.draws == so.draws
    assert list(sg.tries) == list(so.tries)
    assert list(sg.accepts) == list(so.accepts)
    assert sg.bead_updates == so.bead_updates
    assert sg.node_errors == so.node_errors
    L = desc.nlevel
    nm = max(desc.sample[k].nmoving for k in range(desc.nsample))
    plain = sum(sg.tries[l] * (nm << (L - 1 - l)) for l in range(L))
    if desc.sample[0].delayed_rejection:
        assert sg.bead_updates > plain       # second proposals happened
    assert np.max(np.abs(b.get_paths(0) - oc.get_beads())) < 1e-9


def test_delayed_rejection_inputs_are_checked():
    """Delayed rejection needs the Free mover (FreeMoverPBC never sets the
    forward probability the second stage reads); sigma factors must be
    finite and not negative."""
    d = _descs()["pbc"]
    d.sample[0].delayed_rejection = 1
    with pytest.raises(capi.PQError):
        capi.Batch([d])
    d = capi.dot3d_2e_desc()
    d.sample[0].sigma_factor = -1.0
    with pytest.raises(capi.PQError):
        capi.Batch([d])
    d = capi.dot3d_2e_desc()
    d.sample[1].new_sigma_factor = float("nan")
    with pytest.raises(capi.PQError):
        capi.Batch([d])


def test_delayed_rejection_samples_the_exact_bead_law():
    """The device's delayed rejection on its own: 256 chains of one SHO
    particle (2D, m = omega = 1, M = 4, tau = 1/4, L = 1, defaultSigmaFactor
    1.5, newSigmaFactor 3), each on its own Philox stream.  With the exact
    link action each bead's law given its neighbours is Gaussian about
    (x_{t-1} + x_{t+1}) / (2 cosh tau) with variance tanh(tau)/2 per
    coordinate; the chains' mean squared residual matches it."""
    print("[test] delayed rejection bead law on the device", flush=True)
    M, tau, nch = 4, 0.25, 256
    opts = {"delayed_rejection": 1, "sigma_factor": 1.5, "new_sigma_factor": 3.0}
    d = capi.make_desc(2, M, 1, tau, [(1, 1.0, -1.0)], cell=50.0, sho=[(0, 1.0)],
                       samples=[(0, 1, capi.PQ_CHOOSER_SPECIES, 1, opts)], nsection=M)
    b = capi.Batch([d] * nch)
    for c in range(nch):
        b.set_paths(c, np.zeros((M, 1, 2)))
    b.seed(2024, 0)
    b.run_sweeps(100)
    acc = np.zeros(nch)
    nsnap = 150
    for _ in range(nsnap):
        b.run_sweeps(2)
        b.sync()
        for c in range(nch):
            x = b.get_paths(c)[:, 0, :]
            r = x - (np.roll(x, 1, 0) + np.roll(x, -1, 0)) / (2 * np.cosh(tau))
            acc[c] += np.mean(r * r)
    per_chain = acc / nsnap
    exact = np.tanh(tau) / 2
    err = per_chain.std(ddof=1) / np.sqrt(nch)
    assert abs(per_chain.mean() - exact) < max(4 * err, 0.01 * exact), (per_chain.mean(), err, exact)
    st = b.stats(0)
    assert st.bead_updates > st.tries[0]          # second proposals happened
