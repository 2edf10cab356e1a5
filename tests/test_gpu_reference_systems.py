"""The reference's own system tests run on the device through the C ABI.

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
