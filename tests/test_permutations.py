"""Permutation sampling in the oracle, pinned before the device is compared
with it: the reference's Permutation known answers
(unit-tests/util/PermutationTest.cc), the composition rules as functions, the
choosers' draw order on injected streams, the pair action of two moving
particles against a telescoping sum of one-particle moves, and the exchange
statistics of two ideal bosons in a trap against the exact canonical value."""
import math

import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi


def _compose(p, q, append):
    p = np.ascontiguousarray(p, dtype=np.int32).copy()
    q = np.ascontiguousarray(q, dtype=np.int32)
    ol.load().oq_perm_compose(ol.iptr(p), ol.iptr(q), len(p), 1 if append else 0)
    return p


def test_permutation_prepend_kat():
    # PermutationTest.cc:63-80: p = (0 2 3 4 1), its inverse prepended with p
    # is the identity
    p = np.array([0, 2, 3, 4, 1])
    inv = np.empty_like(p)
    inv[p] = np.arange(5)
    assert list(inv) == [0, 4, 1, 2, 3]
    assert list(_compose(inv, p, append=False)) == [0, 1, 2, 3, 4]


def test_permutation_compose_as_functions():
    """prepend(q): p(i) <- p(q(i)); append(q): p(i) <- q(p(i))
    (Permutation.cc:23-43).  Checked on random permutations, where the two
    orders differ."""
    rng = np.random.default_rng(0)
    for _ in range(50):
        p, q = rng.permutation(7), rng.permutation(7)
        assert list(_compose(p, q, append=False)) == [p[q[i]] for i in range(7)]
        assert list(_compose(p, q, append=True)) == [q[p[i]] for i in range(7)]


def test_set_permutation_validates():
    ch = ol.Chain(capi.make_desc(2, 16, 2, 0.1, [(3, 1.0, 0.0)]))
    ch.set_permutation([2, 0, 1])
    assert list(ch.get_permutation()) == [2, 0, 1]
    for bad in ([0, 0, 1], [0, 1, 3], [-1, 0, 1]):
        with pytest.raises(ValueError):
            ch.set_permutation(bad)


def test_energy_wraps_through_the_permutation():
    """The link into slice 0 comes from slice M-1 of inversePermutation[i]
    (SerialPaths::operator(), SerialPaths.cc:68-77).  Two free particles on
    straight paths: under the exchange, particle 0's first bead follows
    particle 1's last bead, and the spring utau of that link says so."""
    M, tau = 8, 0.1
    d = capi.make_desc(1 + 1, M, 1, tau, [(2, 1.0, 0.0)], spring=True, cell=1000.0)
    ch = ol.Chain(d)
    b = np.zeros((M, 2, 2))
    for t in range(M):
        b[t, 0] = [t / M, 0.0]          # particle 0 walks from 0 towards 1
        b[t, 1] = [1.0 + t / M, 0.0]    # particle 1 walks from 1 towards 2
    ch.set_beads(b)
    lam = 0.5
    # identity: slice 0 of particle 0 follows slice M-1 of particle 0
    want_id = 2 / (2 * tau) - (7 / 8) ** 2 / (4 * lam * tau * tau)
    assert ch.bead_utau(0, 0, mask=1) == pytest.approx(want_id, rel=1e-14)
    # particle 1 at slice M-1 continues into particle 0 at slice 0:
    # perm[1] = 0, perm[0] = 1, so inversePermutation[0] = 1
    ch.set_permutation([1, 0])
    want_x = 2 / (2 * tau) - (1 + 7 / 8) ** 2 / (4 * lam * tau * tau)
    assert ch.bead_utau(0, 0, mask=1) == pytest.approx(want_x, rel=1e-14)
    # interior links do not see it
    assert ch.bead_utau(0, 3, mask=1) == pytest.approx(2 / (2 * tau) - (1 / 8) ** 2 / (4 * lam * tau * tau), rel=1e-14)


def _walking_desc(npart, nsize, M=16, L=2, cell=20.0):
    return capi.make_desc(2, M, L, 0.05, [(npart, 1.0, 0.0)], sho=[(0, 1.0)], cell=cell,
                          samples=[(0, nsize, capi.PQ_CHOOSER_WALKING)], nsection=1)


def test_walking_chooser_repeat_ends_the_sample():
    """WalkingChooser.cc:56-64: the second particle is drawn from t(i0,.)/h;
    when it repeats i0 the move ends before the accept draw.  Injected
    stream: section start, i0 = (int)(3 * 0.1) = 0, x = h * 0 < cump(0,0) -> 0."""
    ch = ol.Chain(_walking_desc(3, 2))
    rng = np.random.default_rng(1)
    ch.set_beads(np.repeat(rng.normal(size=(1, 3, 2)), 16, axis=0))
    ch.rng_inject(np.array([0.5, 0.1, 0.0] + [0.5] * 100))
    ch.run_sweeps(1)
    assert ch.lib.oq_rng_count(ch.h) == 3
    st = ch.stats()
    assert st.tries[2] == 0


def test_walking_chooser_cycle_draws_and_moves():
    """A walk that closes goes on to the accept draw (u = 0 < accept) and the
    levels; both particles' first beads stay, and an accepted 2-cycle leaves
    the section's end beads exchanged between them."""
    M, L = 16, 2
    ch = ol.Chain(_walking_desc(2, 2, M=M, L=L))
    rng = np.random.default_rng(2)
    b = np.repeat(rng.normal(size=(1, 2, 2)) * 0.3, M, axis=0)
    ch.set_beads(b)
    # section start 0 (t0 = -1 + (int)(17 u (1-1e-8)) = 0 for u = 1/17 + eps),
    # i0 = 0, x = h * 0.999999 -> particle 1, accept (u = 0), then every level accepted
    u = [1.5 / 17, 0.1, 0.999999, 0.0] + [1e-9] * 400
    ch.rng_inject(np.array(u))
    ch.run_sweeps(1)
    st = ch.stats()
    assert st.tries[L] == 1
    assert st.accepts[0] == 1
    # putBeads relabels the slices after the section and the paths now close
    # on the exchange (SerialPaths.cc:166-184)
    assert list(ch.get_permutation()) == [1, 0]
    after = ch.get_beads()
    S = (1 << L) + 1
    assert np.array_equal(after[0], b[0])
    assert np.array_equal(after[S:, 0], b[S:, 1]) and np.array_equal(after[S:, 1], b[S:, 0])


def test_two_moving_particles_telescope():
    """The action difference of two moving particles (PairAction.cc:241-247
    counts their mutual pair once, with both new paths) equals moving one,
    then the other from the new configuration: the level's action is a
    function of the configuration, so the differences telescope."""
    d = capi.make_desc(3, 64, 3, 0.05, [(2, 1.0, -1.0), (2, 1.0, -1.0)], sho=[(0, 1.0), (1, 1.0)],
                       coulomb=dict(epsilon=1.0, norder=3, rmin=0.005, rmax=50.0, ngpts=1024))
    ch = ol.Chain(d)
    rng = np.random.default_rng(5)
    S = 9
    for level in range(4):
        for i, j in ((0, 1), (0, 2), (3, 1)):
            sec = rng.standard_normal((S, 4, 3))
            mov = np.stack([sec[:, i], sec[:, j]], axis=1).copy()
            mov[1:-1] += 0.3 * rng.standard_normal((S - 2, 2, 3))
            both = ch.action_difference(sec, mov, [i, j], level)
            first = ch.action_difference(sec, mov[:, 0:1].copy(), [i], level)
            mid = sec.copy()
            mid[:, i] = mov[:, 0]
            second = ch.action_difference(mid, mov[:, 1:2].copy(), [j], level)
            assert both == pytest.approx(first + second, rel=1e-12, abs=1e-12), (level, i, j)


def _z1(beta, D):
    return (1.0 / (2 * math.sinh(0.5 * beta))) ** D


def test_two_bosons_exchange_fraction():
    """Two ideal bosons in a 2D trap (omega = 1; spring + SHOAction is the
    exact link propagator): the paths close on the exchange with probability
    Z1(2 beta) / (Z1(beta)^2 + Z1(2 beta)).  One long chain with WalkingChooser
    cycles of two; blocked error over blocks longer than the exchange
    autocorrelation (about a thousand sweeps)."""
    beta, M, L, D = 2.0, 64, 3, 2
    exact = _z1(2 * beta, D) / (_z1(beta, D) ** 2 + _z1(2 * beta, D))
    samples = [(0, 1, capi.PQ_CHOOSER_WALKING), (0, 2, capi.PQ_CHOOSER_WALKING)]
    d = capi.make_desc(D, M, L, beta / M, [(2, 1.0, 0.0)], sho=[(0, 1.0)], samples=samples, cell=20.0)
    ch = ol.Chain(d)
    rng = np.random.default_rng(5)
    ch.set_beads(np.repeat(rng.normal(size=(1, 2, D)), M, axis=0))
    ch.rng_philox(3, 0)
    ch.run_sweeps(1000)
    n = 100000
    x = np.zeros(n)
    for k in range(n):
        ch.run_sweeps(1)
        x[k] = ch.get_permutation()[0] == 1
    blocks = x.reshape(20, -1).mean(1)
    err = blocks.std(ddof=1) / math.sqrt(len(blocks))
    assert abs(x.mean() - exact) < 3.5 * err + 0.005, (x.mean(), exact, err)
