"""Scalar estimator outputs on the host side: the pimc.dat layout of the
reference's ASCII report (src/stats/ascii/AsciiReportBuilder.cc:28-43,
AsciiScalarReportWriter.cpp:12-35) written by the C ABI, and the oracle's
CoulombEnergyEstimator (src/estimator/CoulombEnergyEstimator.cc:30-40)
against closed forms and a brute-force sum with the SuperCell image rule."""
import math
import os

import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi

lib_built = pytest.mark.skipif(not os.path.exists(capi.LIB_PATH), reason="library not built")


@lib_built
def test_ascii_report_layout(tmp_path):
    """Header: '#' and each estimator quoted with its unit in parentheses,
    each followed by a space; rows: (value + shift) * scale at the stream's
    default precision (6 significant digits), each followed by a space."""
    fn = str(tmp_path / "pimc.dat")
    values = np.array([[1.5, -0.25, 2.0], [3.14159265358979, 1e-05, 123456789.0]])
    capi.write_ascii_report(fn, ["thermo_energy", "coulomb_energy", "virial_energy"], values,
                            units=["Ha", "", "meV"], shift=[0.0, 0.0, 1.0], scale=[1.0, 1.0, 0.5])
    want = ('#"thermo_energy (Ha)" "coulomb_energy" "virial_energy (meV)" \n'
            "1.5 -0.25 1.5 \n"
            "3.14159 1e-05 6.17284e+07 \n")
    assert open(fn).read() == want


@lib_built
def test_ascii_report_refuses_missing_names(tmp_path):
    lib = capi.load_library()
    import ctypes as C
    names = (C.c_char_p * 1)(None)
    assert lib.pq_write_ascii_report(str(tmp_path / "x").encode(), 1, names, None, None, None, 0, None) \
        == capi.PQ_ERR_ARG


def test_coulomb_energy_two_charges():
    """Two opposite charges a fixed distance apart in every slice: q1 q2/(eps r)."""
    M = 12
    d = capi.make_desc(3, M, 2, 0.1, [(1, 1.0, -1.0), (1, 2.0, 2.0)], cell=1000.0)
    ch = ol.Chain(d)
    b = np.zeros((M, 2, 3))
    b[:, 1] = [0.3, -0.4, 1.2]
    ch.set_beads(b)
    r = math.sqrt(0.09 + 0.16 + 1.44)
    assert ch.coulomb_energy(1.0) == pytest.approx(-2.0 / r, rel=1e-15)
    assert ch.coulomb_energy(12.0) == pytest.approx(-2.0 / (12.0 * r), rel=1e-15)
    assert ch.coulomb_energy(0.0) == ch.coulomb_energy(1.0)


def _pbc(v, a):
    """SuperCell::pbc (src/util/SuperCell.cc:25-42): minimum image only beyond
    half the shortest side."""
    rcut2 = (0.5 * min(a)) ** 2
    if np.dot(v, v) > rcut2:
        v = v - a * (np.floor(v / a + 1000.5) - 1000)
    return v


def test_coulomb_energy_brute_force_with_images():
    rng = np.random.default_rng(4)
    M, D = 8, 3
    cell = np.array([3.0, 3.5, 4.0])
    species = [(3, 1.0, -1.0), (2, 2.0, 1.0)]
    d = capi.make_desc(D, M, 2, 0.05, species, cell=list(cell))
    ch = ol.Chain(d)
    b = rng.uniform(-1.6, 1.6, size=(M, 5, D))
    ch.set_beads(b)
    q = np.array([-1.0] * 3 + [1.0] * 2)
    want = 0.0
    for t in range(M):
        for i in range(5):
            for j in range(i):
                dv = _pbc(b[t, i] - b[t, j], cell)
                want += q[i] * q[j] / (2.0 * math.sqrt(np.dot(dv, dv)))
    assert ch.coulomb_energy(2.0) == pytest.approx(want / M, rel=1e-13)


def _link_total(ch, t, spring_lam_tau):
    """U_t: the composite link action into slice t summed over particles
    (getBeadAction's u: SHO and half of each pair) plus the spring link action
    |r(t) - r(t-1)|^2 / (4 lambda tau), which SpringAction's getBeadAction
    leaves out of u."""
    b = ch.get_beads()
    M = b.shape[0]
    tot = 0.0
    for k in range(b.shape[1]):
        u, _, _, _ = ch.bead_action(k, t % M)
        tot += u
        if spring_lam_tau:
            prev = b[(t - 1) % M, k]
            tot += np.dot(b[t % M, k] - prev, b[t % M, k] - prev) / (4 * spring_lam_tau)
    return tot


def test_bead_forces_are_link_action_gradients():
    """getBeadAction's fm = -dU_t/dr_i(t) (the link into slice t) and
    fp = -dU_{t+1}/dr_i(t) (the link out of it), for spring + SHO + Coulomb
    pair action, by central differences (SpringAction.cc:112-123,
    SHOAction.cc:71-105, PairAction.cc:509-599 with uk0CalcDerivatives)."""
    d = capi.make_desc(3, 16, 2, 0.05, [(2, 1.0, -1.0), (1, 2.0, 1.0)], sho=[(0, 1.0), (1, 0.6)],
                       coulomb=dict(epsilon=1.0, norder=3, rmin=0.005, rmax=50.0, ngpts=1024))
    ch = ol.Chain(d)
    rng = np.random.default_rng(9)
    base = rng.standard_normal((1, 3, 3)) * 1.2
    b = base + 0.2 * rng.standard_normal((16, 3, 3))
    ch.set_beads(b)
    h = 1e-6
    for i, t in ((0, 5), (1, 9), (2, 3)):
        _, _, fm, fp = ch.bead_action(i, t)
        for dd in range(3):
            got = []
            for sgn in (1, -1):
                bb = b.copy()
                bb[t, i, dd] += sgn * h
                ch.set_beads(bb)
                got.append((_link_total(ch, t, 0.0), _link_total(ch, t + 1, 0.0)))
            ch.set_beads(b)
            # the spring part in closed form of the same link actions
            lam = 0.5 / (1.0 if i < 2 else 2.0)
            sm = -(b[t, i, dd] - b[t - 1, i, dd]) / (2 * lam * d.tau)
            sp = -(b[t, i, dd] - b[t + 1, i, dd]) / (2 * lam * d.tau)
            want_m = -(got[0][0] - got[1][0]) / (2 * h) + sm
            want_p = -(got[0][1] - got[1][1]) / (2 * h) + sp
            assert fm[dd] == pytest.approx(want_m, rel=2e-6, abs=2e-5), (i, t, dd)
            assert fp[dd] == pytest.approx(want_p, rel=2e-6, abs=2e-5), (i, t, dd)


def test_virial_with_unit_window_is_the_thermodynamic_estimator():
    """With nwindow = 1 every link's window closes on itself: the -F.r/2tau
    term and the window's closing F.r0/2tau cancel, leaving sum utau / M."""
    for d in (capi.dot3d_2e_desc(), capi.make_desc(2, 24, 2, 0.05, [(3, 1.0, -1.0)], sho=[(0, 1.0)], cell=3.0,
                                                coulomb=dict(epsilon=2.0, norder=2, ngpts=300))):
        ch = ol.Chain(d)
        ch.set_beads(capi.initial_beads(d, 5, replicate=False))
        ch.rng_philox(3, 0)
        ch.run_sweeps(3)
        assert ch.virial_energy(1) == pytest.approx(ch.energy(), rel=1e-12)
        assert ch.virial_energy(4) != pytest.approx(ch.energy(), rel=1e-6)
