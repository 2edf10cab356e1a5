"""The reference's own system tests run on the device through the C ABI:
test/system/sho here, test/system/hatom below (section groups), and
test/system/madelung in test_gpu_ewald.py.

test/system/sho (pimc.xml, test_sho.py:27-32): one particle in a 3D
oscillator, omega 0.5 Ha, T 1 Ha, 32 slices, <ChooseSection nlevel="5">, so
a section spans the whole path (2^L = M); 100 thermalising sections, then
500 blocks of 200 x (section, Measure).  The reference expects the thermal
energy 0.75 coth(0.25) within 0.08 with an error of 0.041 within 0.01.  The
device's decisions equal the oracle's on the same uniforms, so a Philox seed
the oracle passes with is a fixed, repeatable run; every block mean is also
compared with the oracle's."""
import math

import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi
from gpu_helpers import glibc_uniforms

pytestmark = pytest.mark.gpu


def sho_desc():
    # <Sample npart="1" mover="Free" species="e"/>: species given, no
    # noPermutation attribute, so a WalkingChooser (PIMCParser.cc:364-453)
    return capi.make_desc(3, 32, 5, 1 / 32, [(1, 1.0, -1.0)], cell=20.0, sho=[(0, 0.5)],
                          samples=[(0, 1, capi.PQ_CHOOSER_WALKING)], nsection=1)


def test_whole_path_sections_match_the_oracle():
    print("[test] sho system: decisions on srand(0) uniforms", flush=True)
    d = sho_desc()
    b = capi.Batch([d])
    oc = ol.Chain(d)
    z = np.zeros((32, 1, 3))
    b.set_paths(0, z)
    oc.set_beads(z)
    u = glibc_uniforms(0, 400000)      # <RandomGenerator/> without iseed: srand(0)
    b.set_uniforms(u[None, :])
    oc.rng_inject(u)
    for _ in range(4):
        b.run_sweeps(100)
        b.sync()
        oc.run_sweeps(100)
        assert np.max(np.abs(b.get_paths(0) - oc.get_beads())) < 1e-9
    sg, so = b.stats(0), oc.stats()
    assert sg.draws == so.draws and list(sg.accepts) == list(so.accepts)
    assert abs(b.energy()[0] - oc.energy()) <= 1e-12 * max(1.0, abs(oc.energy()))


def test_reference_sho_system_test():
    print("[test] sho system: 500 blocks of 200 measured sections", flush=True)
    d = sho_desc()
    b = capi.Batch([d])
    oc = ol.Chain(d)
    z = np.zeros((32, 1, 3))
    b.set_paths(0, z)
    oc.set_beads(z)
    b.seed(0, 0)
    oc.rng_philox(0, 0)
    b.run_sweeps(100)
    oc.run_sweeps(100)
    blocks = np.array([b.run_measured(200, 1)[:, 0].mean() for _ in range(500)])
    want = oc.energy_blocks(500, 200)
    np.testing.assert_allclose(blocks, want, rtol=1e-12, atol=1e-12)
    e = blocks.mean()
    de = math.sqrt(blocks.var() / (len(blocks) - 1))   # pitools Scalar.getAverage
    assert abs(e - 0.75 / math.tanh(0.25)) < 0.08, (e, de)
    assert abs(de - 0.041) < 0.01, de


def test_section_groups_match_the_oracle():
    """Sweeps of three <ChooseSection> groups with their own nlevel and
    Samples (the hydrogen-atom input) make the oracle's decisions."""
    from reference_systems import hatom_beads, hatom_desc
    print("[test] section groups: hydrogen atom sweeps against the oracle", flush=True)
    d = hatom_desc(False)
    b = capi.Batch([d])
    oc = ol.Chain(d)
    b.set_paths(0, hatom_beads())
    oc.set_beads(hatom_beads())
    b.seed(3, 0)
    oc.rng_philox(3, 0)
    for _ in range(4):
        b.run_sweeps(5)
        b.sync()
        oc.run_sweeps(5)
        assert np.max(np.abs(b.get_paths(0) - oc.get_beads())) < 1e-9
    sg, so = b.stats(0), oc.stats()
    assert sg.draws == so.draws and list(sg.tries) == list(so.tries) and list(sg.accepts) == list(so.accepts)
    assert sg.tries[7] > 0 and sg.tries[5] > sg.tries[7]


def test_reference_hydrogen_atom_system_test():
    """test/system/hatom (test_hatom.py): 10000 thermalising sections, 100
    blocks of 300 measured sweeps; thermal, virial (nwindow 500) and Coulomb
    energies in eV against the reference's expectations, and every block
    against the oracle's run of the same streams."""
    from reference_systems import EV, HATOM_KT_EV, block_average, hatom_beads, hatom_desc
    from test_oracle_systems import run_hatom_oracle
    print("[test] hydrogen atom system test on the device", flush=True)
    th = capi.Batch([hatom_desc(True)])
    th.set_paths(0, hatom_beads())
    th.seed(7, 0)
    th.run_sweeps(10000)
    b = capi.Batch([hatom_desc(False)])
    b.set_paths(0, th.get_paths(0))
    b.seed(8, 0)
    E, V, C = [], [], []
    for blk in range(100):
        e = v = c = 0.0
        for _ in range(300):
            b.run_sweeps(1)
            e += b.energy()[0]
            v += b.virial_energy(500)[0]
            c += b.coulomb_energy(1.0)[0]
        E.append(e / 300 * EV)
        V.append(v / 300 * EV)
        C.append(c / 300 * EV)
        if blk % 20 == 19:
            print(f"[test] hydrogen atom block {blk + 1} of 100", flush=True)
    print("[test] hydrogen atom device run done; oracle run", flush=True)
    Eo, Vo, Co = run_hatom_oracle()
    np.testing.assert_allclose(E, Eo, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(V, Vo, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(C, Co, rtol=1e-9, atol=1e-9)
    e, de = block_average(E)
    assert abs(e - (-13.6 + 1.5 * HATOM_KT_EV)) < 1.2 and abs(de - 0.60) < 0.2, (e, de)
    v, dv = block_average(V)
    assert abs(v - (-13.6 + 1.5 * HATOM_KT_EV)) < 0.3 and abs(dv - 0.16) < 0.05, (v, dv)
    c, dc = block_average(C)
    assert abs(c - (-2 * 13.6)) < 1.2 and abs(dc - 0.20) < 0.1, (c, dc)


def _group_descs():
    from gpu_helpers import small_dot2d_3p3
    W = capi.PQ_CHOOSER_WALKING
    # fermions with SHO nodes: sections of 2^4 links over both species, then
    # 2 x 2^3 over one species and 3 x 2^2 over the other
    ferm = small_dot2d_3p3(64)
    capi.add_group(ferm, 3, [(0, 1)], 2)
    capi.add_group(ferm, 2, [(1, 1)], 3)
    # bosons in a trap: one-particle sections, then WalkingChooser cycles of
    # two and three particles at other depths
    bos = capi.make_desc(2, 64, 3, 2.0 / 64, [(4, 1.0, 0.0)], sho=[(0, 1.0)], cell=20.0,
                         samples=[(0, 1, W)], nsection=4)
    capi.add_group(bos, 4, [(0, 2, W)], 2)
    capi.add_group(bos, 2, [(0, 3, W), (0, 1, W)], 3)
    # <Sample nrepeat>: three tries per section for the fermions' first
    # Sample, two for a WalkingChooser cycle of the bosons
    rep_f = small_dot2d_3p3(64)
    rep_f.sample[0].nrepeat = 3
    rep_b = capi.make_desc(2, 64, 3, 2.0 / 64, [(4, 1.0, 0.0)], sho=[(0, 1.0)], cell=20.0,
                           samples=[(0, 2, W, 2), (0, 1, W, 3)], nsection=4)
    return {"fermions": ferm, "bosons": bos, "repeat-fermions": rep_f, "repeat-bosons": rep_b}


@pytest.mark.parametrize("name", ["fermions", "bosons", "repeat-fermions", "repeat-bosons"])
def test_section_groups_with_nodes_and_permutations(name):
    print(f"[test] section groups {name}", flush=True)
    d = _group_descs()[name]
    b = capi.Batch([d])
    oc = ol.Chain(d)
    beads = capi.initial_beads(d, 11, replicate=True)
    b.set_paths(0, beads)
    oc.set_beads(beads)
    b.seed(17, 0)
    oc.rng_philox(17, 0)
    for _ in range(4):
        b.run_sweeps(3)
        b.sync()
        oc.run_sweeps(3)
        assert list(b.get_permutation(0)) == list(oc.get_permutation())
        assert np.max(np.abs(b.get_paths(0) - oc.get_beads())) < 1e-9
    sg, so = b.stats(0), oc.stats()
    assert sg.draws == so.draws and list(sg.tries) == list(so.tries) and list(sg.accepts) == list(so.accepts)
    assert sg.node_errors == so.node_errors
    e, eo = b.energy()[0], oc.energy()
    assert abs(e - eo) <= 1e-12 * max(1.0, abs(eo)), (e, eo)
