"""The Measure step on the device against the oracle: the load-balanced
thermodynamic-energy items (every chain shape, permuted paths, periodic
cells, static species), the pair-correlation histogram with a partner list
(bit-exact, periodic images included), and the measured-sweep graph with the
histogram inside it against the same steps run one call at a time, over
several replays."""
import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi
from gpu_helpers import rel_close

pytestmark = pytest.mark.gpu


def _box3():
    # three species in a periodic box small enough that pbc images matter
    return capi.make_desc(3, 48, 3, 0.05, [(5, 1.0, -1.0), (3, 2.0, 1.0), (4, 1.0, -1.0)], cell=[3.0, 3.5, 2.5],
                          coulomb=dict(epsilon=1.0, norder=3, rmin=0.0, rmax=0.0, ngpts=300),
                          mover=capi.PQ_MOVER_FREE_PBC)


def test_pair_cf_partner_list_bit_exact():
    print("[test] pair cf partner list", flush=True)
    d = _box3()
    b = capi.Batch([d, d])
    rng = np.random.default_rng(9)
    ocs = []
    for c in range(2):
        beads = rng.uniform(-2.0, 2.0, size=(d.nslice, d.npart, 3))
        b.set_paths(c, beads)
        oc = ol.Chain(d)
        oc.set_beads(beads)
        ocs.append(oc)
    for species in ([0, 1, 2], [1, 1], [2, 0, 0], [0], [1, 2, 1, 0]):
        got = b.pair_cf_histogram(species, 0.1, 2.0, 64)
        for c, oc in enumerate(ocs):
            want = oc.pair_cf_histogram(species, 0.1, 2.0, 64)
            assert np.array_equal(got[c], want), (species, c)
    # the two-species call is the list of two
    assert np.array_equal(b.pair_histogram(2, 0, 0.1, 2.0, 64), b.pair_cf_histogram([2, 0], 0.1, 2.0, 64))


def test_pair_cf_bin_edges_bit_exact():
    """Pair distances exactly on bin edges, one ulp either side of them, at
    rmin and at rmax: the device's bins equal the oracle's FP64 expression
    (correctly rounded square root, no contraction), for a bin width that is
    a power of two and for ones that are not."""
    print("[test] pair cf bin edges", flush=True)
    n = 24
    d = capi.make_desc(3, 8, 1, 0.05, [(n, 1.0, -1.0)], coulomb=None)
    b = capi.Batch([d])
    oc = ol.Chain(d)
    rng = np.random.default_rng(3)
    for rmin, rmax, nbin in ((0.0, 8.0, 256), (0.1, 2.0, 64), (0.25, 3.0, 37)):
        w = (rmax - rmin) / nbin
        beads = np.zeros((d.nslice, n, 3))
        for t in range(d.nslice):
            ks = rng.integers(0, nbin + 2, size=n)
            x = rmin + ks * w
            x = np.where(rng.random(n) < 0.3, np.nextafter(x, np.inf), x)
            x = np.where(rng.random(n) < 0.3, np.nextafter(x, -np.inf), x)
            # particle 0 at the origin: its distances are the x's; the others
            # meet each other at differences of edges
            x[0] = 0.0
            beads[t, :, 0] = x
        b.set_paths(0, beads)
        oc.set_beads(beads)
        got = b.pair_cf_histogram([0, 0], rmin, rmax, nbin)
        want = oc.pair_cf_histogram([0, 0], rmin, rmax, nbin)
        assert np.array_equal(got[0], want), (rmin, rmax, nbin)
        assert want.sum() > 0


def _shapes():
    out = []
    for n, m in ((2, 16), (7, 200), (20, 1024), (33, 96), (64, 2048), (1, 40)):
        sp = [((n + 1) // 2, 1.0, -1.0)] + ([(n // 2, 1.0, -1.0)] if n > 1 else [])
        out.append(capi.make_desc(3, m, 3, 0.05, sp, sho=[(0, 0.8)],
                                  coulomb=dict(epsilon=1.3, norder=2, rmin=0.005, rmax=50.0, ngpts=700)))
    # a periodic cell with a static species and no SHO
    st = capi.make_desc(3, 64, 3, 0.05, [(4, 1.0, -1.0), (3, 1e9, 1.0, True)], cell=4.0,
                        coulomb=dict(epsilon=1.0, norder=3, rmin=0.0, rmax=0.0, ngpts=400),
                        mover=capi.PQ_MOVER_FREE_PBC)
    out.append(st)
    return out


def test_energy_items_every_shape():
    """Chains of 1..64 particles and 16..2048 slices in one batch, a periodic
    cell with a static species, and a permuted path (slice -1 through the
    inverse permutation): thermodynamic energy against the oracle at 1e-12."""
    print("[test] energy items every shape", flush=True)
    descs = _shapes()
    b = capi.Batch(descs)
    rng = np.random.default_rng(4)
    want = []
    for c, d in enumerate(descs):
        beads = rng.standard_normal((d.nslice, d.npart, 3)) * 1.2
        b.set_paths(c, beads)
        oc = ol.Chain(d)
        oc.set_beads(beads)
        if d.npart >= 7 and d.nspecies == 2 and d.nslice == 200:
            perm = np.arange(d.npart, dtype=np.int32)
            perm[[0, 1, 2]] = perm[[1, 2, 0]]
            b.set_permutation(c, perm)
            oc.set_permutation(perm)
        want.append(oc.energy())
    got = b.energy()
    for c in range(len(descs)):
        assert rel_close(got[c], want[c], 1e-12, scale=10.0), (c, got[c], want[c])


def test_measured_graph_with_pair_cf_matches_eager():
    """nblock x (sweeps, energy, g(r)) in one graph, replayed twice, against
    the same steps launched one by one on a twin batch: identical energies and
    counts."""
    print("[test] measured graph with pair cf", flush=True)
    descs = [d for d in capi.batch2d_descs(16) if d.nslice <= 256][:4]
    twins = []
    for _ in range(2):
        b = capi.Batch(descs)
        for c, d in enumerate(descs):
            b.set_paths(c, capi.initial_beads(d, 70 + c))
        b.seed(5, 0)
        twins.append(b)
    g, e = twins
    g.set_measure_pair_cf([0, 1], 0.0, 8.0, 256)
    for replay in range(2):
        eg = g.run_measured(3, 2)
        cg = g.collect_pair_cf()
        ce = np.zeros_like(cg)
        ee = []
        for blk in range(3):
            e.run_sweeps(2)
            ee.append(e.energy())
            ce += e.pair_cf_histogram([0, 1], 0.0, 8.0, 256)
        assert np.array_equal(eg, np.array(ee)), replay
        assert np.array_equal(cg, ce), replay
        assert cg.sum() > 0
    t_e, ms = g.run_measured_timed(3, 2)
    assert ms.shape == (3,) and np.all(ms > 0)
    for blk in range(3):
        e.run_sweeps(2)
        assert np.array_equal(t_e[blk], e.energy())
    g.set_measure_pair_cf([])
    with pytest.raises(capi.PQError):
        g.collect_pair_cf()
