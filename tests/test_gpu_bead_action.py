"""Action::getBeadAction at the call level (pq_bead_action): u, utau, ulambda
and the forces fm, fp of CompositeAction::getBeadAction on caller-supplied
paths (with a permutation), against the oracle at 1e-12, and the reference's
VirialEnergyEstimator (VirialEnergyEstimator.cc:27-78) driven by it the way a
reference-side estimator shim would, against the device estimator."""
import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi
from gpu_helpers import rel_close

pytestmark = pytest.mark.gpu


def _cases():
    c1 = capi.dot3d_2e_desc()
    box = capi.make_desc(3, 32, 3, 0.05, [(3, 1.0, -1.0), (2, 1e9, 1.0, True)], cell=3.0,
                         coulomb=dict(epsilon=1.0, norder=3, rmin=0.0, rmax=0.0, ngpts=400),
                         mover=capi.PQ_MOVER_FREE_PBC)
    ew = capi.make_desc(3, 16, 2, 0.05, [(2, 1.0, -1.0), (2, 1.0, 1.0)], cell=4.0,
                        coulomb=dict(epsilon=1.0, norder=2, rmin=0.0, rmax=0.0, ngpts=300, ewald=True,
                                     ewald_rcut=2.0, ewald_kcut=6.0),
                        mover=capi.PQ_MOVER_FREE_PBC)
    dot2 = capi.make_desc(2, 40, 3, 0.02, [(3, 1.0, -1.0), (3, 1.0, -1.0)], sho=[(0, 1.0), (1, 1.3)],
                          coulomb=dict(epsilon=2.0, norder=3, rmin=0.005, rmax=50.0, ngpts=1024))
    return {"dot3d": c1, "box": box, "ewald": ew, "dot2": dot2}


@pytest.mark.parametrize("name", ["dot3d", "box", "ewald", "dot2"])
def test_bead_action_on_caller_paths(name):
    print(f"[test] bead action {name}", flush=True)
    d = _cases()[name]
    b = capi.Batch([d])
    rng = np.random.default_rng(21)
    M, N, D = d.nslice, d.npart, d.ndim
    beads = rng.standard_normal((M, N, D)) * (0.8 if d.cell[0] > 100 else 0.6 * d.cell[0])
    perm = np.arange(N, dtype=np.int32)
    if N >= 3:
        perm[[0, 1, 2]] = perm[[2, 0, 1]]
    # the device holds other paths: the call must use the caller's
    b.set_paths(0, np.zeros((M, N, D)))
    oc = ol.Chain(d)
    oc.set_beads(beads)
    oc.set_permutation(perm)
    ii = np.array([i for i in range(N) for t in (0, 1, M // 2, M - 1)], dtype=np.int32)
    tt = np.array([t for i in range(N) for t in (0, 1, M // 2, M - 1)], dtype=np.int32)
    got = b.bead_action(0, ii, tt, beads=beads, perm=perm)
    for q in range(len(ii)):
        u, ut, fm, fp = oc.bead_action(int(ii[q]), int(tt[q]))
        scale = max(1.0, abs(ut))
        assert rel_close(got[q, 0], u, 1e-12, scale=max(1.0, abs(u))), (q, got[q, 0], u)
        assert rel_close(got[q, 1], ut, 1e-12, scale=scale), (q, got[q, 1], ut)
        assert got[q, 2] == 0.0
        fs = max(1.0, float(np.max(np.abs(np.concatenate([fm, fp])))))
        assert rel_close(got[q, 3:3 + D], fm, 1e-12, scale=fs), q
        assert rel_close(got[q, 3 + D:], fp, 1e-12, scale=fs), q
    # the same paths resident on the device give the same values
    b.set_paths(0, beads)
    b.set_permutation(0, perm)
    assert np.array_equal(b.bead_action(0, ii, tt), got)


def _virial_via_bead_action(b, d, beads, nwindow):
    """VirialEnergyEstimator::handleLink/endCalc over SerialPaths::sumOverLinks
    (firstSlice 0), fed by pq_bead_action; identity permutation."""
    M, N, D = d.nslice, d.npart, d.ndim
    tau = d.tau
    ii = np.array([i for t in range(M) for i in range(N)], dtype=np.int32)
    tt = np.array([t for t in range(M) for i in range(N)], dtype=np.int32)
    act = b.bead_action(0, ii, tt, beads=beads).reshape(M, N, -1)
    static = [d.species[s].is_static for s in range(d.nspecies) for _ in range(d.species[s].count)]
    energy = 0.0
    r0 = np.zeros((N, D)); rav = np.zeros((N, D)); fav = np.zeros((N, D))
    for t in range(M):
        for i in range(N):
            if t % nwindow == 0 and not static[i]:
                energy += (1 / (2 * tau)) * fav[i] @ r0[i]
                r0[i] = 0; rav[i] = 0; fav[i] = 0
            W = nwindow if t // nwindow < (M - 1) // nwindow else (M - 1) % nwindow + 1
            r = beads[t, i] - beads[t - 1, i]     # open dot: no images
            rav[i] += r
            r0[i] += (W - t % nwindow) * r / W
            utau, fm, fp = act[t, i, 1], act[t, i, 3:3 + D], act[t, i, 3 + D:]
            fav[i] += fm + fp
            energy += utau
            if not static[i]:
                energy += -(1. - 1. / W) * (D / (2 * tau)) - (1 / (2 * tau)) * (fm + fp) @ rav[i]
    for i in range(N):
        if not static[i]:
            energy += (1 / (2 * tau)) * fav[i] @ r0[i]
    return energy / M


def test_virial_estimator_through_the_call_level():
    print("[test] virial through pq_bead_action", flush=True)
    d = _cases()["dot2"]
    b = capi.Batch([d])
    rng = np.random.default_rng(3)
    base = rng.standard_normal((1, d.npart, 2))
    beads = base + 0.2 * rng.standard_normal((d.nslice, d.npart, 2))
    b.set_paths(0, beads)
    for nwindow in (1, 4):
        want = float(b.virial_energy(nwindow)[0])
        got = _virial_via_bead_action(b, d, beads, nwindow)
        assert rel_close(got, want, 1e-11, scale=10.0), (nwindow, got, want)
