"""Device parity of the Ewald long-range Coulomb sum (CoulombAction useEwald,
OptEwaldSum, EwaldCoulombEstimator) and of static species, through the C ABI
against the oracle, which tests/test_ewald.py pins to the reference's Madelung
test and to an independent error-function Ewald sum."""
import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi
from gpu_helpers import rel_close
from test_ewald import MADELUNG, nacl_beads, nacl_desc

pytestmark = pytest.mark.gpu


def box_desc(D=3, L=3.0, nslice=32, nlevel=3, eps=1.0, kcut=None, nsection=None, ewald_type="opt"):
    """Two +1 and two -1 charges (masses 1 and 2) in a periodic box with the
    Ewald sum (optEwald, or tradEwald with the parser's default kappa),
    FreeMoverPBC, one Sample per species."""
    kc = kcut if kcut is not None else 2 * np.pi * 3 / L
    return capi.make_desc(D, nslice, nlevel, 0.05, [(2, 1.0, 1.0), (2, 2.0, -1.0)], cell=L,
                          coulomb=dict(epsilon=eps, norder=2, ngpts=400, rmin=0.005, rmax=L * np.sqrt(D),
                                       ewald=True, ewald_rcut=L / 2, ewald_kcut=kc, ewald_type=ewald_type),
                          mover=capi.PQ_MOVER_FREE_PBC, nsection=nsection)


def box_beads(desc, seed):
    rs = np.random.default_rng(seed)
    L = desc.cell[0]
    base = rs.uniform(0, L, size=(1, desc.npart, desc.ndim))
    b = base + 0.1 * rs.standard_normal((desc.nslice, desc.npart, desc.ndim))
    return np.ascontiguousarray(b)


def test_madelung_on_device():
    print("[test] ewald madelung", flush=True)
    d = nacl_desc()
    b = capi.Batch([d])
    b.set_paths(0, nacl_beads())
    oc = ol.Chain(d)
    oc.set_beads(nacl_beads())
    e = b.energy()[0]
    assert rel_close(e, oc.energy(), 1e-12, scale=1.0), (e, oc.energy())
    assert abs(e / 4 - MADELUNG) < 1e-4
    ec = b.ewald_coulomb_energy(1.0, 0.5, 50.0)[0]
    assert rel_close(ec, oc.ewald_coulomb_energy(1.0, 0.5, 50.0), 1e-12, scale=1.0), ec
    assert abs(ec / 4 - MADELUNG) < 1e-5
    ev = b.virial_energy(1)[0]
    assert rel_close(ev, oc.virial_energy(1), 1e-11, scale=1.0), (ev, oc.virial_energy(1))


@pytest.mark.parametrize("D,eps,prec,kind", [(3, 1.0, capi.PQ_PREC_FP64, "opt"), (2, 2.0, capi.PQ_PREC_FP64, "opt"),
                                             (3, 1.0, capi.PQ_PREC_MIXED, "opt"), (3, 1.0, capi.PQ_PREC_FP64, "trad")])
def test_ewald_action_difference(D, eps, prec, kind):
    print(f"[test] ewald action difference D={D} eps={eps} prec={prec} {kind}", flush=True)
    d = box_desc(D=D, eps=eps, L=3.0 if D == 3 else 4.0, kcut=None if D == 3 else 10.0, ewald_type=kind)
    b = capi.Batch([d], precision=prec)
    oc = ol.Chain(d)
    rng = np.random.default_rng(21 + D)
    S = (1 << d.nlevel) + 1
    L = d.cell[0]
    secs, movs, idxs, lev, want = [], [], [], [], []
    for q in range(24):
        sec = rng.uniform(0, L, size=(1, d.npart, D)) + 0.15 * rng.standard_normal((S, d.npart, D))
        i = int(rng.integers(0, d.npart))
        mov = sec[:, [i], :].copy()
        mov[1:-1] += 0.2 * rng.standard_normal((S - 2, 1, D))
        level = 0 if q % 2 == 0 else q % (d.nlevel + 1)
        secs.append(sec); movs.append(mov); idxs.append(i); lev.append(level)
        want.append(oc.action_difference(sec, mov, [i], level))
    nq = len(secs)
    got = np.zeros(nq)
    sec_all = np.ascontiguousarray(np.concatenate([s.ravel() for s in secs]))
    mov_all = np.ascontiguousarray(np.concatenate([m.ravel() for m in movs]))
    capi.check(b.lib.pq_action_difference(b.h, nq, capi.iptr(np.zeros(nq, dtype=np.int32)),
                                          capi.iptr(np.array(lev, dtype=np.int32)), capi.dptr(sec_all),
                                          capi.dptr(mov_all), capi.iptr(np.array(idxs, dtype=np.int32)), 1,
                                          capi.dptr(got)))
    tol = 1e-12 if prec == capi.PQ_PREC_FP64 else 5e-5
    assert rel_close(got, np.array(want), tol, scale=1.0), np.max(np.abs(got - np.array(want)))
    # the level-0 queries carry a nonzero Ewald term
    ew = [oc.action_difference(secs[q], movs[q], [idxs[q]], 0, mask=8) for q in range(0, nq, 2)]
    assert min(abs(x) for x in ew) > 1e-8


@pytest.mark.parametrize("team,kind", [(-1, "opt"), (128, "opt"), (-1, "trad")])
def test_ewald_sampler_decisions(team, kind):
    print(f"[test] ewald sampler team={team} {kind}", flush=True)
    d = box_desc(nslice=32, nlevel=3, ewald_type=kind)
    b = capi.Batch([d])
    b.set_schedule(team, -1)
    oc = ol.Chain(d)
    beads = box_beads(d, 7)
    b.set_paths(0, beads)
    oc.set_beads(beads)
    b.seed(31, 0)
    oc.rng_philox(31, 0)
    for _ in range(3):
        b.run_sweeps(2)
        b.sync()
        oc.run_sweeps(2)
        assert np.max(np.abs(b.get_paths(0) - oc.get_beads())) < 1e-9
    sg, so = b.stats(0), oc.stats()
    assert sg.draws == so.draws
    assert list(sg.tries) == list(so.tries)
    assert list(sg.accepts) == list(so.accepts)
    assert sg.accepts[0] > 0
    e, eo = b.energy()[0], oc.energy()
    assert rel_close(e, eo, 1e-12, scale=1.0), (e, eo)
    ev, evo = b.virial_energy(2)[0], oc.virial_energy(2)
    assert rel_close(ev, evo, 1e-11, scale=1.0), (ev, evo)
    ec, eco = b.ewald_coulomb_energy(1.0, 1.5, 8.0)[0], oc.ewald_coulomb_energy(1.0, 1.5, 8.0)
    assert rel_close(ec, eco, 1e-12, scale=1.0), (ec, eco)


def test_bead_utau_carries_the_long_range_part():
    print("[test] ewald bead utau", flush=True)
    d = box_desc(nslice=16, nlevel=2)
    b = capi.Batch([d])
    oc = ol.Chain(d)
    beads = box_beads(d, 3)
    b.set_paths(0, beads)
    oc.set_beads(beads)
    qi = np.array([0, 1, 2, 3, 0, 2], dtype=np.int32)
    qt = np.array([0, 5, 9, 15, 7, 3], dtype=np.int32)
    out = np.zeros(len(qi))
    capi.check(b.lib.pq_bead_utau(b.h, len(qi), capi.iptr(np.zeros(len(qi), dtype=np.int32)), capi.iptr(qi),
                                  capi.iptr(qt), capi.dptr(out)))
    want = np.array([oc.bead_utau(int(i), int(t)) for i, t in zip(qi, qt)])
    assert rel_close(out, want, 1e-12, scale=1.0), (out, want)


def test_bad_ewald_inputs_are_refused():
    d = box_desc()
    d.coulomb.ewald_kcut = 0.0
    with pytest.raises(capi.PQError):
        capi.Batch([d])
    b = capi.Batch([box_desc()])
    with pytest.raises(capi.PQError):
        b.ewald_coulomb_energy(1.0, 0.0, 5.0)


def test_trad_ewald_estimator_on_device():
    """pq_measure_trad_ewald_coulomb_energy: NaCl's Madelung energy with
    kappa 5, and the oracle's values on random boxes (3D and 2D, the parser's
    default kappa and an explicit one, two cells in one batch); switching
    between the optEwald and tradEwald estimators rebuilds the cached fits."""
    print("[test] trad ewald estimator", flush=True)
    d = nacl_desc()
    b = capi.Batch([d])
    b.set_paths(0, nacl_beads())
    oc = ol.Chain(d)
    oc.set_beads(nacl_beads())
    et = b.ewald_coulomb_energy(1.0, 0.5, 50.0, trad=True, kappa=5.0)[0]
    assert rel_close(et, oc.trad_ewald_coulomb_energy(1.0, 0.5, 50.0, 5.0), 1e-12, scale=1.0), et
    assert abs(et / 4 - MADELUNG) < 1e-8
    eo = b.ewald_coulomb_energy(1.0, 0.5, 50.0)[0]
    assert rel_close(eo, oc.ewald_coulomb_energy(1.0, 0.5, 50.0), 1e-12, scale=1.0), eo
    for D in (3, 2):
        descs = [box_desc(D=D, L=3.0), box_desc(D=D, L=4.0)]
        b = capi.Batch(descs)
        ocs = []
        for c, dd in enumerate(descs):
            beads = capi.initial_beads(dd, 60 + c)
            beads = np.mod(beads, dd.cell[0])
            b.set_paths(c, beads)
            o = ol.Chain(dd)
            o.set_beads(beads)
            ocs.append(o)
        for kappa in (0.0, 2.5):
            for eps in (1.0, 2.0):
                got = b.ewald_coulomb_energy(eps, 1.4, 6.0, trad=True, kappa=kappa)
                for c, o in enumerate(ocs):
                    want = o.trad_ewald_coulomb_energy(eps, 1.4, 6.0, kappa)
                    assert rel_close(got[c], want, 1e-12, scale=1.0), (D, kappa, eps, c, got[c], want)
    with pytest.raises(capi.PQError):
        b.ewald_coulomb_energy(1.0, 1.4, 6.0, trad=True, kappa=-1.0)
