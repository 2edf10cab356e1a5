"""Inputs of the reference's system tests (test/system/*/pimc.xml) as chain
descriptions, with the units of src/stats/Units.cc."""
import numpy as np

import piqmc_capi as capi

AMU = 9.1093826e-31 / 1.66053886e-27   # electron masses per amu (Units.cc:44)
EV = 27.2113845                        # eV per Hartree (Units.cc:25)
KELVIN = 3.1577465e5                   # K per Hartree (Units.cc:29)
ANGSTROM = 0.5291772086                # A per bohr (Units.cc:15)

# test/system/hatom/pimc.xml: an electron and a proton, T = 10000 K,
# 512 slices, a = 25 A, CoulombAction norder 3 rmin 0.001 rmax 5 with 1000
# grid points; one Sample (WalkingChooser, Free mover) per species
HATOM_M = 512
HATOM_T = 10000 / KELVIN
HATOM_KT_EV = 0.0316682 * 27.211       # test_hatom.py:19


def hatom_desc(thermalize):
    """thermalize: <Loop nrepeat="10000"> of <ChooseSection nlevel="6"> over
    both species; else the measured sweep: nlevel 7 over both species, then
    3 x nlevel 6 and 5 x nlevel 5 over the electron, before each Measure."""
    sp = [(1, 1.0, -1.0), (1, 1.006780 / AMU, 1.0)]
    W = capi.PQ_CHOOSER_WALKING
    co = dict(norder=3, rmin=0.001, rmax=5.0, ngpts=1000)
    tau = 1 / (HATOM_T * HATOM_M)
    cell = 25 / ANGSTROM
    if thermalize:
        return capi.make_desc(3, HATOM_M, 6, tau, sp, cell=cell, coulomb=co,
                              samples=[(1, 1, W), (0, 1, W)], nsection=1)
    d = capi.make_desc(3, HATOM_M, 7, tau, sp, cell=cell, coulomb=co,
                       samples=[(1, 1, W), (0, 1, W)], nsection=1)
    capi.add_group(d, 6, [(0, 1, W)], 3)
    capi.add_group(d, 5, [(0, 1, W)], 5)
    return d


def hatom_beads():
    """The proton at the origin and the electron one bohr away on every
    slice (the reference's SetCubicLattice spreads the two particles over a
    2 x 2 x 2 lattice of side 1 bohr)."""
    b = np.zeros((HATOM_M, 2, 3))
    b[:, 0, 0] = 1.0
    return b


def block_average(x):
    """pitools Scalar.getAverage: mean and sqrt(var / (n - 1))."""
    x = np.asarray(x)
    return float(x.mean()), float(np.sqrt(x.var() / (len(x) - 1)))
