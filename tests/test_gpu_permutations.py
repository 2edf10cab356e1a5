"""GPU parity of permutation sampling: Samples that move several particles,
WalkingChooser cycles, the permuted getBeads/putBeads bookkeeping and the
path permutation, through the C ABI against the oracle on the same uniforms.
Decisions and permutations are identical, paths agree to 1e-9 after several
sweeps, FP64 action differences to 1e-12."""
import math

import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi
from gpu_helpers import glibc_uniforms, rel_close, small_dot2d_3p3

pytestmark = pytest.mark.gpu

W, SP, ALL, ALLR = (capi.PQ_CHOOSER_WALKING, capi.PQ_CHOOSER_SPECIES, capi.PQ_CHOOSER_ALL,
                    capi.PQ_CHOOSER_ALL_RANDOM)


def _perm_descs():
    trap = dict(sho=[(0, 1.0)], cell=20.0)
    bosons = capi.make_desc(2, 64, 3, 2.0 / 64, [(4, 1.0, 0.0)], samples=[(0, 1, W), (0, 2, W), (0, 3, W)],
                            **trap)
    # a periodic box small enough that the walking table, spring and mover use
    # the theta series (FreeMoverPBC)
    box = capi.make_desc(3, 32, 2, 0.1, [(3, 1.0, 0.0)], cell=2.5, mover=capi.PQ_MOVER_FREE_PBC,
                         samples=[(0, 2, W), (0, 1, W)])
    coul = dict(epsilon=1.0, norder=3, rmin=0.005, rmax=50.0, ngpts=1024)
    mixed = capi.make_desc(3, 64, 3, 0.05, [(3, 1.0, -1.0), (2, 2.0, -1.0)], sho=[(0, 1.0), (1, 1.0)],
                           coulomb=coul, samples=[(0, 2, ALLR), (0, 3, ALL), (1, 2, SP)])
    # fermions: cycles of 2 cross the nodes at the top level, cycles of 3 may not;
    # DoubleSectionChooser's putDoubleBeads relabels section 2
    ferm = small_dot2d_3p3(64)
    ferm.nsample = 2
    ferm.sample[0].species, ferm.sample[0].nmoving, ferm.sample[0].chooser = 0, 3, W
    ferm.sample[1].species, ferm.sample[1].nmoving, ferm.sample[1].chooser = 1, 2, W
    return {"bosons": bosons, "box": box, "mixed": mixed, "fermions": ferm}


def _run_pair(desc, nsweep, chunks, inject, team=-1, seed=5):
    b = capi.Batch([desc])
    b.set_schedule(team, -1)
    oc = ol.Chain(desc)
    beads = capi.initial_beads(desc, seed, replicate=True)
    b.set_paths(0, beads)
    oc.set_beads(beads)
    if inject:
        u = glibc_uniforms(seed, 2000000)
        b.set_uniforms(u[None, :])
        oc.rng_inject(u)
    else:
        b.seed(99, 0)
        oc.rng_philox(99, 0)
    perms = []
    for _ in range(chunks):
        b.run_sweeps(nsweep)
        b.sync()
        oc.run_sweeps(nsweep)
        pg, po = b.get_permutation(0), oc.get_permutation()
        assert list(pg) == list(po)
        perms.append(tuple(pg))
        assert np.max(np.abs(b.get_paths(0) - oc.get_beads())) < 1e-9
    return b, oc, perms


@pytest.mark.parametrize("name,inject,team", [("bosons", False, -1), ("bosons", True, 64), ("box", False, 128),
                                              ("mixed", True, -1), ("mixed", False, 64), ("fermions", False, -1),
                                              ("fermions", True, 64)])
def test_permutation_sampler_decisions(name, inject, team):
    print(f"[test] permutation sampler {name} inject={inject} team={team}", flush=True)
    desc = _perm_descs()[name]
    b, oc, perms = _run_pair(desc, 4, 5, inject, team)
    sg, so = b.stats(0), oc.stats()
    assert sg.draws == so.draws
    assert list(sg.tries) == list(so.tries)
    assert list(sg.accepts) == list(so.accepts)
    assert sg.bead_updates == so.bead_updates
    assert sg.accepted_bead_updates == so.accepted_bead_updates
    assert sg.node_errors == so.node_errors
    if name in ("bosons", "box"):
        # exchanges were accepted along the way
        assert any(p != tuple(range(desc.npart)) for p in perms)
    # the energy estimator follows the permutation across slice 0
    e = b.energy()[0]
    assert rel_close(e, oc.energy(), 1e-12, scale=1.0), (e, oc.energy())


def test_several_moving_batches_run_small_teams():
    desc = _perm_descs()["bosons"]
    b = capi.Batch([desc, desc])
    b.set_schedule(256, -1)
    assert b.schedule()[0] == 128
    assert b.lib.pq_set_split(b.h, 1, 256) == capi.PQ_ERR_ARG


def test_action_difference_two_moving():
    """pq_action_difference with nmov = 2: the moving pair counted once with
    both new paths (PairAction.cc:241-247), same-species and cross-species."""
    print("[test] action_difference two moving", flush=True)
    desc = capi.dot3d_2e_desc()
    desc2 = capi.make_desc(3, 64, 3, 0.05, [(2, 1.0, -1.0), (2, 2.0, 1.0)], sho=[(0, 1.0), (1, 0.7)],
                           coulomb=dict(epsilon=1.0, norder=3, rmin=0.005, rmax=50.0, ngpts=1024))
    for d in (desc, desc2):
        b = capi.Batch([d])
        oc = ol.Chain(d)
        rng = np.random.default_rng(13)
        S = (1 << d.nlevel) + 1
        secs, movs, idxs, lev, want = [], [], [], [], []
        for q in range(40):
            sec = rng.standard_normal((1, d.npart, d.ndim)) + 0.3 * rng.standard_normal((S, d.npart, d.ndim))
            pair = rng.choice(d.npart, size=2, replace=False)
            mov = np.stack([sec[:, pair[0]], sec[:, pair[1]]], axis=1).copy()
            mov[1:-1] += 0.3 * rng.standard_normal((S - 2, 2, d.ndim))
            if q % 3 == 0:
                # a 2-cycle: the moving particles end on each other's last bead
                mov[-1, 0], mov[-1, 1] = sec[-1, pair[1]], sec[-1, pair[0]]
            level = q % (d.nlevel + 1)
            secs.append(sec); movs.append(mov); idxs.extend(int(p) for p in pair); lev.append(level)
            want.append(oc.action_difference(sec, mov, [int(p) for p in pair], level))
        nq = len(secs)
        got = np.zeros(nq)
        sec_all = np.ascontiguousarray(np.concatenate([s.ravel() for s in secs]))
        mov_all = np.ascontiguousarray(np.concatenate([m.ravel() for m in movs]))
        ch = np.zeros(nq, dtype=np.int32)
        lv = np.array(lev, dtype=np.int32)
        ix = np.array(idxs, dtype=np.int32)
        capi.check(b.lib.pq_action_difference(b.h, nq, capi.iptr(ch), capi.iptr(lv), capi.dptr(sec_all),
                                              capi.dptr(mov_all), capi.iptr(ix), 2, capi.dptr(got)))
        assert rel_close(got, np.array(want), 1e-12, scale=1.0), np.max(np.abs(got - np.array(want)))


def test_double_action_difference_two_moving():
    """FixedNodeAction with two moving particles of one species, a 2-cycle
    included: the top level sees the exchanged last slice (a sign change, so
    2e100), and the rest agrees with the oracle to 1e-11."""
    print("[test] double action two moving", flush=True)
    desc = small_dot2d_3p3(64)
    b = capi.Batch([desc])
    oc = ol.Chain(desc)
    rng = np.random.default_rng(21)
    S = (1 << desc.nlevel) + 1
    base = rng.standard_normal((1, desc.npart, 2))
    sec1 = np.ascontiguousarray(base + 0.15 * rng.standard_normal((S, desc.npart, 2)))
    sec2 = np.ascontiguousarray(base + 0.15 * rng.standard_normal((S, desc.npart, 2)))
    capi.check(b.lib.pq_section_begin(b.h, 0, capi.dptr(sec1), capi.dptr(sec2)))
    oc.fn_initialize(sec1, sec2)
    nbig = 0
    for trial in range(12):
        pair = [int(v) for v in rng.choice(3, size=2, replace=False)]
        mov = np.ascontiguousarray(np.stack([sec1[:, pair[0]], sec1[:, pair[1]]], axis=1).copy())
        mov[1:-1] += 0.05 * rng.standard_normal((S - 2, 2, 2))
        if trial % 3 == 2:
            mov[-1, 0], mov[-1, 1] = sec1[-1, pair[1]].copy(), sec1[-1, pair[0]].copy()
        ix = np.array(pair, dtype=np.int32)
        for level in range(desc.nlevel, -1, -1):
            got = np.zeros(1)
            capi.check(b.lib.pq_double_action_difference(b.h, 0, level, capi.dptr(mov), capi.iptr(ix), 2,
                                                         capi.dptr(got)))
            want = oc.fn_action_difference(mov, pair, level)
            if want >= 1e100:
                nbig += 1
                assert got[0] == want
                break
            assert rel_close(got[0], want, 1e-11, scale=1.0), (trial, level, got[0], want)
        else:
            capi.check(b.lib.pq_accept_last_move(b.h, 0))
            oc.fn_accept()
            sec1[:, pair[0]] = mov[:, 0]
            sec1[:, pair[1]] = mov[:, 1]
    assert nbig >= 1


def test_energy_with_a_set_permutation():
    desc = capi.dot3d_2e_desc()
    desc3 = capi.make_desc(2, 64, 3, 0.05, [(3, 1.0, -1.0), (2, 1.0, -1.0)], sho=[(0, 1.0), (1, 1.0)],
                           coulomb=dict(epsilon=1.0, norder=3, rmin=0.005, rmax=50.0, ngpts=1024))
    rng = np.random.default_rng(2)
    for d in (desc, desc3):
        b = capi.Batch([d])
        oc = ol.Chain(d)
        beads = capi.initial_beads(d, 3, replicate=False)
        b.set_paths(0, beads)
        oc.set_beads(beads)
        p = rng.permutation(d.npart).astype(np.int32)
        b.set_permutation(0, p)
        oc.set_permutation(p)
        assert list(b.get_permutation(0)) == list(p)
        assert rel_close(b.energy()[0], oc.energy(), 1e-12, scale=1.0)
        with pytest.raises(capi.PQError):
            b.set_permutation(0, np.zeros(d.npart, dtype=np.int32))
        # new paths start from the identity again
        b.set_paths(0, beads)
        assert list(b.get_permutation(0)) == list(range(d.npart))


def test_batch_path_files_keep_the_permutation(tmp_path):
    desc = _perm_descs()["bosons"]
    b, oc, perms = _run_pair(desc, 6, 3, False)
    base = str(tmp_path / "bpaths")
    capi.check(b.lib.pq_write_all_paths(b.h, base.encode()))
    b2 = capi.Batch([desc])
    capi.check(b2.lib.pq_read_all_paths(b2.h, base.encode(), 1))
    assert list(b2.get_permutation(0)) == list(b.get_permutation(0))
    assert np.allclose(b2.get_paths(0), b.get_paths(0), rtol=1e-14, atol=1e-14)
    assert rel_close(b2.energy()[0], b.energy()[0], 1e-12, scale=1.0)


def _z1(beta, D):
    return (1.0 / (2 * math.sinh(0.5 * beta))) ** D


def test_sampled_boson_exchange_fraction():
    """1024 chains of two ideal bosons in a 2D trap: the fraction of paths
    closing on the exchange against Z1(2 beta)/(Z1(beta)^2 + Z1(2 beta)).
    Four snapshots 1000 sweeps apart after 3000 sweeps of equilibration."""
    print("[test] boson exchange fraction", flush=True)
    beta, M, L, D = 2.0, 64, 3, 2
    exact = _z1(2 * beta, D) / (_z1(beta, D) ** 2 + _z1(2 * beta, D))
    d = capi.make_desc(D, M, L, beta / M, [(2, 1.0, 0.0)], sho=[(0, 1.0)], cell=20.0,
                       samples=[(0, 1, W), (0, 2, W)])
    nchain = 1024
    b = capi.Batch([d] * nchain)
    rng = np.random.default_rng(8)
    for c in range(nchain):
        b.set_paths(c, np.repeat(rng.normal(size=(1, 2, D)), M, axis=0))
    b.seed(17, 0)
    b.run_sweeps(3000)
    hits = []
    for snap in range(4):
        b.run_sweeps(1000)
        b.sync()
        hits.append(np.mean([b.get_permutation(c)[0] == 1 for c in range(nchain)]))
    frac = float(np.mean(hits))
    err = math.sqrt(exact * (1 - exact) / (4 * nchain))
    print(f"[test] exchange fraction {frac:.4f} exact {exact:.4f} binomial err {err:.4f}", flush=True)
    assert abs(frac - exact) < 5 * err
