"""Oracle checks of the Ewald long-range Coulomb sum (CoulombAction useEwald,
OptEwaldSum, EwaldCoulombEstimator), on the CPU.

Pins:
* the reference's own Madelung system test (test/system/madelung/pimc.xml,
  test_nacl.py): NaCl in a cubic cell a = 2 with static ions, Coulomb energy
  per ion pair -1.747564594633 within 1e-5 and thermal energy within 1e-4;
* the fit's constraints: f(rcut) = 1/rcut and f'(rcut) = -1/rcut^2
  (OptEwaldSum.cc:136-166);
* the Ewald identity against an independent error-function Ewald sum
  (numpy/scipy) on random neutral configurations in 3D and 2D;
* the level-0 action term tau/eps sum_t [LR(new) - LR(old)] over slices
  1..2^L-1 against the same independent sum, and the pair tables' u_0 and
  utau_0 columns shifted by -tau q1q2 f(r) and -q1q2 f(r).
"""
import itertools
import math

import numpy as np
import pytest
from scipy.special import erfc

import oracle_lib as ol
import piqmc_capi as capi

AMU = 9.1093826e-31 / 1.66053886e-27      # Units.cc:44, electron masses per amu
MADELUNG = -1.747564594633182190636      # test/system/madelung/test_nacl.py


def nacl_desc(M=10):
    # test/system/madelung/pimc.xml: a=2, Na 23 amu +1, Cl 35 amu -1, both
    # static; T = 1000 Ha, nslice 10; CoulombAction norder 1, rmin 0.01,
    # rmax 3, 1000 grid points, Ewald rcut 0.5, kcut 50
    return capi.make_desc(3, M, 1, 1.0 / (1000.0 * M),
                          [(4, 23 / AMU, 1.0, 1), (4, 35 / AMU, -1.0, 1)], cell=2.0,
                          coulomb=dict(norder=1, rmin=0.01, rmax=3.0, ngpts=1000, ewald=True,
                                       ewald_rcut=0.5, ewald_kcut=50.0),
                          samples=[(0, 1)], nsection=1)


def nacl_beads(M=10):
    # test/system/madelung/paths.in: Na at the fcc sites, Cl on the edges
    na = [(0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
    cl = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)]
    return np.tile(np.array(na + cl, dtype=float)[None], (M, 1, 1))


def test_madelung_energies():
    ch = ol.Chain(nacl_desc())
    ch.set_beads(nacl_beads())
    e_coul = ch.ewald_coulomb_energy(1.0, 0.5, 50.0) / 4      # perN="4"
    assert abs(e_coul - MADELUNG) < 1e-5, e_coul
    e_thermo = ch.energy() / 4
    assert abs(e_thermo - MADELUNG) < 1e-4, e_thermo


def test_trad_ewald_estimator_madelung():
    """<CoulombEnergyEstimator useEwald="true" ewaldType="tradEwald">: the
    TradEwaldSum estimator gives the NaCl Madelung energy when erfc(kappa r)
    has died out by the nearest neighbour (r = 1, kappa = 5) and the k-space
    Gaussian by kcut (exp(-kcut^2 / 4 kappa^2) = e^-25); the parser's default
    kappa is sqrt((3.6 N)^(1/6) sqrt(pi) / Lside)."""
    ch = ol.Chain(nacl_desc())
    ch.set_beads(nacl_beads())
    e = ch.trad_ewald_coulomb_energy(1.0, 0.5, 50.0, 5.0) / 4
    assert abs(e - MADELUNG) < 1e-8, e
    assert ch.trad_estimator_kappa() == pytest.approx(math.sqrt((3.6 * 8) ** (1 / 6) * math.sqrt(math.pi) / 2.0),
                                                      rel=1e-15)
    # with the default kappa (1.4) the real-space tail beyond rcut is lost
    assert abs(ch.trad_ewald_coulomb_energy(1.0, 0.5, 50.0) / 4 - MADELUNG) > 1e-3


def test_fit_meets_its_constraints():
    for D, L, rc, kc in [(3, 2.0, 0.5, 50.0), (3, 5.0, 2.5, 8.0), (2, 4.0, 2.0, 10.0)]:
        d = capi.make_desc(D, 8, 1, 0.1, [(2, 1.0, 1.0), (2, 1.0, -1.0)], cell=L,
                           coulomb=dict(norder=0, ngpts=200, ewald=True, ewald_rcut=rc,
                                        ewald_kcut=kc))
        ch = ol.Chain(d)
        coef = ch.ewald_info()[0]
        f = sum(c * rc ** (2 * i) for i, c in enumerate(coef))
        df = sum(2 * i * c * rc ** (2 * i - 1) for i, c in enumerate(coef))
        # the 8 x 8 normal equations are ill-conditioned: the constraints hold
        # to rounding of the solve, a few parts in 1e9
        assert abs(f - 1 / rc) < 1e-7 / rc, (D, f, 1 / rc)
        assert abs(df + 1 / rc ** 2) < 1e-6 / rc ** 2, (D, df, -1 / rc ** 2)
        assert abs(ch.ewald_fr(rc) - 1 / rc) < 1e-7 / rc
        assert ch.ewald_fr(1.3 * rc) == 1 / (1.3 * rc)


def erfc_ewald(pos, q, L, D, kappa):
    """Periodic Coulomb energy of a neutral cell by the error-function Ewald
    sum: real space over the nearest images, k space over |n| <= nmax."""
    N = len(q)
    e = 0.0
    shifts = list(itertools.product((-1, 0, 1), repeat=D))
    for i in range(N):
        for j in range(N):
            for s in shifts:
                if i == j and not any(s):
                    continue
                r = np.linalg.norm(pos[i] - pos[j] + L * np.array(s))
                e += 0.5 * q[i] * q[j] * erfc(kappa * r) / r
    nmax = int(math.ceil(2 * kappa * 6.5 * L / (2 * math.pi)))
    rng = np.arange(-nmax, nmax + 1)
    ks = np.array(list(itertools.product(rng, repeat=D)), dtype=float) * (2 * math.pi / L)
    ks = ks[np.any(ks != 0, axis=1)]
    k = np.linalg.norm(ks, axis=1)
    rho = (q[None, :] * np.exp(1j * ks @ pos.T)).sum(axis=1)
    if D == 3:
        w = 4 * math.pi / k ** 2 * np.exp(-k ** 2 / (4 * kappa ** 2))
    else:
        w = 2 * math.pi / k * erfc(k / (2 * kappa))
    e += 0.5 / L ** D * float(np.sum(w * np.abs(rho) ** 2))
    e -= kappa / math.sqrt(math.pi) * float(np.sum(q ** 2))
    return e


def short_range(ch, pos, q, L, rc):
    """sum_{i<j} q_i q_j (1/r - f(r)) over minimum images inside rcut."""
    e = 0.0
    for i in range(len(q)):
        for j in range(i):
            dv = pos[i] - pos[j]
            dv -= L * np.round(dv / L)
            r = float(np.linalg.norm(dv))
            if r < rc:
                e += q[i] * q[j] * (1 / r - ch.ewald_fr(r))
    return e


# The 2D fit sums the Bessel series of OptEwaldSum.cc:92-118 by terms up to
# (k rcut/2)^2m/m!^2 in the halo k < 4 kcut (skipped only above k rcut = 30):
# at kcut rcut = 20 its cancellation leaves the fit accurate to 1e-6 of the
# energy, by kcut rcut = 40 it is NaN, in the reference as here.
@pytest.mark.parametrize("D,L,rc,kc", [(3, 3.0, 1.5, 14.0), (2, 4.0, 2.0, 10.0)])
def test_long_range_matches_erfc_ewald(D, L, rc, kc):
    rs = np.random.default_rng(11 + D)
    d = capi.make_desc(D, 8, 1, 0.1, [(3, 1.0, 1.0), (3, 1.0, -1.0)], cell=L,
                       coulomb=dict(norder=0, ngpts=200, ewald=True, ewald_rcut=rc, ewald_kcut=kc))
    ch = ol.Chain(d)
    q = np.array([1.0] * 3 + [-1.0] * 3)
    kappa = 5.0 / (L / 2)
    for _ in range(3):
        pos = rs.uniform(0, L, size=(6, D))
        exact = erfc_ewald(pos, q, L, D, kappa)
        opt = short_range(ch, pos, q, L, rc) + ch.ewald_long_range(pos)
        assert abs(opt - exact) < 2e-5 * max(1.0, abs(exact)), (D, opt, exact)


def test_level0_ewald_term_and_tables():
    D, L, rc, kc, eps, tau = 3, 3.0, 1.5, 14.0, 2.0, 0.05
    rs = np.random.default_rng(5)
    coul = dict(epsilon=eps, norder=2, ngpts=300, rmin=0.01, rmax=6.0, ewald=True,
                ewald_rcut=rc, ewald_kcut=kc)
    d = capi.make_desc(D, 16, 2, tau, [(2, 1.0, 1.0), (2, 2.0, -1.0)], cell=L, coulomb=coul)
    ch = ol.Chain(d)
    q = np.array([1.0, 1.0, -1.0, -1.0])
    S = 5
    sec = rs.uniform(0, L, size=(S, 4, D))
    i = 2
    mov = sec[:, [i], :].copy()
    mov[1:S - 1] += rs.normal(0, 0.2, size=(S - 2, 1, D))
    kappa = 5.0 / (L / 2)

    def lr(pos):
        return erfc_ewald(pos, q, L, D, kappa) - short_range(ch, pos, q, L, rc)

    want = 0.0
    for t in range(1, S - 1):
        new = sec[t].copy()
        new[i] = mov[t, 0]
        want += (lr(new) - lr(sec[t])) * tau / eps
    got = ch.action_difference(sec, mov, [i], 0, mask=8)
    assert abs(got - want) < 2e-5 * max(1.0, abs(want)), (got, want)
    assert ch.action_difference(sec, mov, [i], 1, mask=8) == 0.0
    # CoulombAction whole = pair tables + the Ewald term
    full = ch.action_difference(sec, mov, [i], 0, mask=4)
    d0 = capi.make_desc(D, 16, 2, tau, [(2, 1.0, 1.0), (2, 2.0, -1.0)], cell=L,
                        coulomb={k: v for k, v in coul.items() if k != "ewald"})
    ch0 = ol.Chain(d0)
    assert ch.ewald_info() is not None and ch0.ewald_info() is None
    # tables: u_0 loses tau q1q2 f(r), utau_0 loses q1q2 f(r), higher orders unchanged
    for ip in range(3):
        meta, grid, u, ut = ch.pair_table(ip)
        meta0, grid0, u0, ut0 = ch0.pair_table(ip)
        s1, s2 = int(meta[0]), int(meta[1])
        q1q2 = [1.0, -1.0][s1] * [1.0, -1.0][s2] / eps
        r = (1 / grid[0]) * np.exp(np.arange(int(meta[3])) / grid[1])
        fr = np.array([ch.ewald_fr(x) for x in r])
        np.testing.assert_allclose(u[:, 0], u0[:, 0] - tau * q1q2 * fr, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(ut[:, 0], ut0[:, 0] - q1q2 * fr, rtol=1e-6, atol=1e-8)
        np.testing.assert_array_equal(u[:, 1:], u0[:, 1:])
    assert np.isfinite(full)


def test_static_species_skip_the_spring():
    # a static particle adds no spring action and no kinetic utau
    d = capi.make_desc(3, 8, 2, 0.1, [(1, 1.0, 0.0, 1), (1, 1.0, 0.0)], cell=1000.0)
    ch = ol.Chain(d)
    rs = np.random.default_rng(3)
    b = rs.normal(size=(8, 2, 3))
    ch.set_beads(b)
    assert ch.bead_utau(0, 3, 1) == 0.0
    assert ch.bead_utau(1, 3, 1) != 0.0
    S = 5
    sec = b[:S]
    mov = sec[:, [0], :] + 0.1
    assert ch.action_difference(sec, mov, [0], 0, mask=1) == 0.0
    assert ch.action_difference(sec, sec[:, [1], :] + 0.1, [1], 0, mask=1) != 0.0


def test_reference_erf():
    from scipy.special import erf
    lib = ol.load()
    xs = np.concatenate([np.linspace(-6, 6, 241), [1e-8, 0.5, 1.2247, 1.2248, 3.0]])
    got = np.array([lib.oq_ref_erf(float(x)) for x in xs])
    np.testing.assert_allclose(got, erf(xs), rtol=2e-15, atol=2e-16)
    # the reference's own erf.cpp, compiled into oracle/_ref: bit for bit
    ref = ol.load_ref()
    if ref is None:
        pytest.skip("reference sources not mounted: oracle/_ref not built")
    rs = np.random.default_rng(4)
    for x in np.concatenate([xs, rs.uniform(-8, 8, 2000), rs.uniform(-1.3, 1.3, 500)]):
        assert lib.oq_ref_erf(float(x)) == ref.ref_erf(float(x)), x


def test_traditional_ewald_matches_erfc_ewald():
    """ewaldType="tradEwald": the tables lose tau q1q2 erf(kappa r)/r at
    every r, and the k-space sum with 4 PI exp(-k^2/4 kappa^2)/k^2 restores
    it; the total equals the error-function Ewald sum with the same kappa."""
    D, L = 3, 3.0
    kappa = 5.0 / (L / 2)
    d = capi.make_desc(D, 8, 1, 0.1, [(3, 1.0, 1.0), (3, 1.0, -1.0)], cell=L,
                       coulomb=dict(norder=0, ngpts=200, ewald=True, ewald_type="trad",
                                    ewald_kappa=kappa, ewald_kcut=40.0))
    ch = ol.Chain(d)
    q = np.array([1.0] * 3 + [-1.0] * 3)
    rs = np.random.default_rng(17)
    for _ in range(3):
        pos = rs.uniform(0, L, size=(6, D))
        exact = erfc_ewald(pos, q, L, D, kappa)
        sr = 0.0
        for i in range(6):
            for j in range(i):
                dv = pos[i] - pos[j]
                dv -= L * np.round(dv / L)
                r = float(np.linalg.norm(dv))
                sr += q[i] * q[j] * (1 / r - ch.ewald_fr(r))
        assert abs(sr + ch.ewald_long_range(pos) - exact) < 1e-10 * max(1.0, abs(exact))
    # the parser's default kappa, sqrt((3.6 N)^(1/6) sqrt(pi) / min a)
    d.coulomb.ewald_kappa = 0.0
    ch0 = ol.Chain(d)
    k0 = math.sqrt((3.6 * 6) ** (1 / 6) * math.sqrt(math.pi) / L)
    assert ch0.ewald_fr(0.7) == pytest.approx(ol.load().oq_ref_erf(k0 * 0.7) / 0.7, rel=1e-15)
