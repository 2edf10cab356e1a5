"""The reference's hydrogen-atom system test (test/system/hatom/pimc.xml,
test_hatom.py) on the oracle: 10000 thermalising sections, then 100 blocks of
300 sweeps of three section groups (nlevel 7 over both particles, 3 x nlevel 6
and 5 x nlevel 5 over the electron), each followed by the thermal, virial
(nwindow 500) and Coulomb estimators; the reference's expectations and
tolerances (in eV)."""
import oracle_lib as ol
from reference_systems import (EV, HATOM_KT_EV, block_average, hatom_beads, hatom_desc)


def run_hatom_oracle(nblock=100, nmeasure=300):
    th = ol.Chain(hatom_desc(True))
    th.set_beads(hatom_beads())
    th.rng_philox(7, 0)
    th.run_sweeps(10000)
    ch = ol.Chain(hatom_desc(False))
    ch.set_beads(th.get_beads())
    ch.rng_philox(8, 0)
    E, V, C = [], [], []
    for _ in range(nblock):
        e = v = c = 0.0
        for _ in range(nmeasure):
            ch.run_sweeps(1)
            e += ch.energy()
            v += ch.virial_energy(500)
            c += ch.coulomb_energy(1.0)
        E.append(e / nmeasure * EV)
        V.append(v / nmeasure * EV)
        C.append(c / nmeasure * EV)
    return E, V, C


def test_hydrogen_atom_system_test():
    E, V, C = run_hatom_oracle()
    e, de = block_average(E)
    assert abs(e - (-13.6 + 1.5 * HATOM_KT_EV)) < 1.2 and abs(de - 0.60) < 0.2, (e, de)
    v, dv = block_average(V)
    assert abs(v - (-13.6 + 1.5 * HATOM_KT_EV)) < 0.3 and abs(dv - 0.16) < 0.05, (v, dv)
    c, dc = block_average(C)
    assert abs(c - (-2 * 13.6)) < 1.2 and abs(dc - 0.20) < 0.1, (c, dc)
