"""The pi-qmc input reader (pq_read_xml / pq_parse_xml) on inputs these tests
write themselves: the structure of the reference's SHO and hydrogen-atom
system tests (test/system/{sho,hatom}/pimc.xml) gives the same descriptors
as tests/reference_systems.py builds by hand, units are scaled with the
reference's table (Units.cc:9-48), fermion inputs map node models, samplers
and the pair-correlation estimator, and every element or attribute the device
path does not model is refused with a message that names it."""
import ctypes as C

import pytest

import piqmc_capi as capi
import reference_systems as rs


def _fields_equal(a, b, path="desc"):
    """Field-by-field equality of two ctypes structures (arrays included)."""
    for name, typ in a._fields_:
        va, vb = getattr(a, name), getattr(b, name)
        if isinstance(va, C.Structure):
            _fields_equal(va, vb, f"{path}.{name}")
        elif isinstance(va, C.Array):
            for k in range(len(va)):
                if isinstance(va[k], C.Structure):
                    _fields_equal(va[k], vb[k], f"{path}.{name}[{k}]")
                else:
                    assert va[k] == pytest.approx(vb[k], rel=1e-15, abs=0), f"{path}.{name}[{k}]"
        else:
            assert va == pytest.approx(vb, rel=1e-15, abs=0), f"{path}.{name}"


SHO = """<?xml version="1.0"?>
<!-- one particle in a 3D oscillator: the layout of the reference's SHO system test -->
<Simulation>
  <SuperCell a="20.0" x="1" y="1" z="1"/>
  <Species name="e" count="1" mass="1" charge="-1"/>
  <Temperature value="1.0" nslice="32"/>
  <Action>
    <SpringAction/>
    <SHOAction omega="0.5 Ha"/>
  </Action>
  <Estimators><ThermalEnergyEstimator/></Estimators>
  <PIMC>
    <RandomGenerator/>
    <Loop nrepeat="100">
      <ChooseSection nlevel="5"><Sample npart="1" mover="Free" species="e"/></ChooseSection>
    </Loop>
    <Loop nrepeat="500">
      <Loop nrepeat="200">
        <ChooseSection nlevel="5"><Sample npart="1" mover="Free" species="e"/></ChooseSection>
        <Measure estimator="all"/>
      </Loop>
      <Collect estimator="all"/>
    </Loop>
  </PIMC>
</Simulation>
"""


def test_sho_input_gives_the_system_test_descriptor():
    x = capi.read_xml(SHO, 3, text=True)
    want = capi.make_desc(3, 32, 5, 1 / 32, [(1, 1.0, -1.0)], cell=20.0, sho=[(0, 0.5)],
                          samples=[(0, 1, capi.PQ_CHOOSER_WALKING)], nsection=1)
    _fields_equal(x.sweep, want)
    assert (x.nmeasure, x.nblock, x.nthermalize, x.has_thermalize) == (200, 500, 100, 1)
    _fields_equal(x.thermalize, want)
    assert x.has_seed == 1 and x.iseed == 0
    assert x.thermal_energy == 1 and x.virial_nwindow == 0 and x.coulomb_energy == 0


HATOM = """<Simulation>
  <SuperCell a="25 A" x="1" y="1" z="1"/>
  <Species name="eup" count="1" mass="1 m_e" charge="-1"/>
  <Species name="H" count="1" mass="1.006780 amu" charge="+1"/>
  <Temperature value="10000 K" nslice="512"/>
  <Action>
    <SpringAction/>
    <CoulombAction norder="3" rmin="0.001" rmax="5." ngridPoints="1000"/>
  </Action>
  <Estimators>
    <ThermalEnergyEstimator unit="eV"/>
    <VirialEnergyEstimator unit="eV" nwindow="500"/>
    <CoulombEnergyEstimator unit="eV"/>
  </Estimators>
  <PIMC>
    <RandomGenerator iseed="17"/>
    <SetCubicLattice nx="2" ny="2" nz="2" a="1."/>
    <Loop nrepeat="10000">
      <ChooseSection nlevel="6">
         <Sample npart="1" mover="Free" species="H"/>
         <Sample npart="1" mover="Free" species="eup"/>
      </ChooseSection>
    </Loop>
    <Loop nrepeat="100">
      <Loop nrepeat="300">
        <ChooseSection nlevel="7">
          <Sample npart="1" mover="Free" species="H"/>
          <Sample npart="1" mover="Free" species="eup"/>
        </ChooseSection>
        <Loop nrepeat="3"><ChooseSection nlevel="6"><Sample npart="1" mover="Free" species="eup"/></ChooseSection></Loop>
        <Loop nrepeat="5"><ChooseSection nlevel="5"><Sample npart="1" mover="Free" species="eup"/></ChooseSection></Loop>
        <Measure estimator="all"/>
      </Loop>
      <Collect estimator="all"/>
    </Loop>
  </PIMC>
</Simulation>"""


def test_hatom_input_with_units_and_section_groups(tmp_path):
    f = tmp_path / "pimc.xml"
    f.write_text(HATOM)
    x = capi.read_xml(str(f), 3)
    _fields_equal(x.sweep, rs.hatom_desc(False))
    _fields_equal(x.thermalize, rs.hatom_desc(True))
    assert (x.nmeasure, x.nblock, x.nthermalize) == (300, 100, 10000)
    assert x.iseed == 17 and x.nignored == 1
    assert x.thermal_energy == 1 and x.virial_nwindow == 500 and x.coulomb_energy == 1
    assert x.coulomb_epsilon == 1.0 and x.coulomb_ewald == 0
    # the unit scales themselves
    assert x.sweep.cell[0] == pytest.approx(25 / 0.5291772086, rel=1e-15)
    assert x.sweep.tau == pytest.approx(1 / (10000 / 3.1577465e5 * 512), rel=1e-15)


FERMIONS = """<Simulation>
  <SuperCell a="10" x="1" y="1"/>
  <Species name="up" count="3" mass="1" charge="-1"/>
  <Species name="dn" count="3" mass="1" charge="-1"/>
  <Species name="h" count="2" mass="2" charge="1" isStatic="True"/>
  <Temperature value="0.125" tau="0.02 Ha-1"/>
  <Action>
    <SpringAction pgDelta="0.02"/>
    <SHOAction omega="1" species="up" x="0.5"/>
    <SHOAction omega="1" species="dn" ndim="1"/>
    <CoulombAction epsilon="2" norder="2" ngridPoints="400" useEwald="true" ewaldRcut="5" ewaldKcut="8" ewaldType="trad"/>
    <FixedNodeAction species="up" model="SHONodes" omega="1" temperature="0.2"/>
    <FixedNodeAction species="dn" noNodalAction="true" useManyBodyDistance="true"/>
  </Action>
  <Estimators>
    <PairCFEstimator name="g_up" nspecies="3" species1="up" species2="up" species3="dn">
      <Radial nbin="64" min="0" max="4"/>
    </PairCFEstimator>
    <CoulombEnergyEstimator useEwald="true" rcut="5" kcut="8" ewaldType="tradEwald" kappa="0.7"/>
  </Estimators>
  <PIMC>
    <ReadPaths file="paths.in" bfactor="2"/>
    <Loop nrepeat="4000">
      <Loop nrepeat="25">
        <ChooseSection nlevel="4">
          <Sample npart="2" species="up" nrepeat="3"/>
          <Sample npart="1" species="dn" noPermutation="true" delayedRejection="true"/>
          <Sample npart="1" species="all"/>
          <Sample npart="2" noPermutation="true" defaultSigmaFactor="0.8" newSigmaFactor="0.6"/>
        </ChooseSection>
      </Loop>
      <Measure estimator="all"/>
    </Loop>
  </PIMC>
</Simulation>"""


def test_fermion_input_maps_nodes_samplers_and_estimators():
    x = capi.read_xml(FERMIONS, 2, text=True)
    d = x.sweep
    assert (d.ndim, d.nspecies, d.nslice) == (2, 3, 400)
    assert d.tau == pytest.approx(0.02)
    assert list(d.cell)[:2] == [10.0, 10.0]
    assert d.species[2].is_static == 1 and d.species[2].mass == 2.0
    assert d.nsho == 2 and d.sho[0].center[0] == 0.5 and d.sho[1].species == 1 and d.sho[1].ndim == 1
    c = d.coulomb
    assert (c.enabled, c.epsilon, c.norder, c.ngpts, c.use_ewald, c.ewald_type) == (1, 2.0, 2, 400, 1, 1)
    assert (c.ewald_rcut, c.ewald_kcut, c.ewald_kappa) == (5.0, 8.0, 0.0)
    assert d.nnodes == 2
    assert (d.nodes[0].model, d.nodes[0].omega, d.nodes[0].temperature, d.nodes[0].nodal_action) == \
        (capi.PQ_NODES_SHO, 1.0, 0.2, 1)
    assert (d.nodes[1].model, d.nodes[1].nodal_action, d.nodes[1].many_body_distance) == (capi.PQ_NODES_FREE, 0, 1)
    # mover="" is FreeMoverPBC (PIMCParser.cc:332-335)
    assert d.mover == capi.PQ_MOVER_FREE_PBC
    assert (d.nlevel, d.nsample, d.nsection, d.ngroup) == (4, 4, 25, 0)
    s = d.sample
    assert (s[0].species, s[0].nmoving, s[0].chooser, s[0].nrepeat) == (0, 2, capi.PQ_CHOOSER_WALKING, 3)
    assert (s[1].species, s[1].chooser, s[1].delayed_rejection) == (1, capi.PQ_CHOOSER_SPECIES, 1)
    assert s[2].chooser == capi.PQ_CHOOSER_ALL_RANDOM
    assert (s[3].chooser, s[3].sigma_factor, s[3].new_sigma_factor) == (capi.PQ_CHOOSER_ALL, 0.8, 0.6)
    assert (x.nmeasure, x.nblock, x.has_thermalize) == (4000, 1, 0)
    assert x.read_paths == b"paths.in" and x.read_bfactor == 2
    assert x.pair_cf_nspecies == 3 and list(x.pair_cf_species)[:3] == [0, 0, 1]
    assert (x.pair_cf_rmin, x.pair_cf_rmax, x.pair_cf_nbin) == (0.0, 4.0, 64)
    assert (x.coulomb_energy, x.coulomb_ewald, x.coulomb_ewald_type, x.coulomb_kappa) == (1, 1, 1, 0.7)


def _refused(text, ndim=3):
    with pytest.raises(capi.PQError) as e:
        capi.read_xml(text, ndim, text=True)
    return str(e.value)


@pytest.mark.parametrize("edit,word", [
    (("<SpringAction/>", "<SpringAction/><GaussianAction v0='1' alpha='2'/>"), "GaussianAction"),
    (('omega="0.5 Ha"', 'omega="0.5 furlong"'), "furlong"),
    (('mover="Free" species="e"/></ChooseSection>\n    </Loop>',
      'mover="FreePBC" species="e"/></ChooseSection>\n    </Loop>'), None),
    (("<ThermalEnergyEstimator/>", "<ThermalEnergyEstimator/><DensityEstimator/>"), "DensityEstimator"),
    (("</Simulation>", ""), "never closed"),
    (('<Loop nrepeat="500">', '<Loop nrepeat="500" x="1" x="2">'), "twice"),
])
def test_inputs_outside_the_device_path_are_refused(edit, word):
    text = SHO.replace(*edit, 1)
    if word is None:
        # a different mover in the thermalising loop than in the sweep is
        # allowed (two descriptors); mixing movers inside one sweep is not
        x = capi.read_xml(text, 3, text=True)
        assert x.thermalize.mover == capi.PQ_MOVER_FREE_PBC and x.sweep.mover == capi.PQ_MOVER_FREE
        mixed = SHO.replace('<Sample npart="1" mover="Free" species="e"/></ChooseSection>\n        <Measure',
                            '<Sample npart="1" mover="Free" species="e"/><Sample npart="1" species="e"/>'
                            '</ChooseSection>\n        <Measure')
        assert "different movers" in _refused(mixed)
        return
    assert word in _refused(text)


@pytest.mark.parametrize("attr", ['useHungarian="true"', 'useIterations="2"'])
def test_free_particle_node_variants_are_refused(attr):
    text = FERMIONS.replace('noNodalAction="true"', f'noNodalAction="true" {attr}')
    assert "FreeParticleNodes" in _refused(text, 2)


def test_use_updates_reads_into_the_node_descriptor():
    """useUpdates="true" on FreeParticleNodes selects FreeParticleNodes::
    MatrixUpdate (ActionParser.cc:808-811); absent it stays off"""
    x = capi.read_xml(FERMIONS, 2, text=True)
    assert [x.sweep.nodes[a].use_updates for a in range(x.sweep.nnodes)] == [0, 0]
    text = FERMIONS.replace('noNodalAction="true"', 'noNodalAction="true" useUpdates="true"')
    x = capi.read_xml(text, 2, text=True)
    assert [x.sweep.nodes[a].model for a in range(2)] == [capi.PQ_NODES_SHO, capi.PQ_NODES_FREE]
    assert [x.sweep.nodes[a].use_updates for a in range(2)] == [0, 1]
    # SHONodes takes no useUpdates
    text = text.replace('temperature="0.2"/>', 'temperature="0.2" useUpdates="true"/>')
    x = capi.read_xml(text, 2, text=True)
    assert [x.sweep.nodes[a].use_updates for a in range(2)] == [0, 1]


def test_coulomb_variants_are_refused():
    assert "exchangeLevel" in _refused(FERMIONS.replace('ewaldType="trad"', 'ewaldType="trad" exchangeLevel="1"'), 2)
    assert "ewaldImages" in _refused(FERMIONS.replace('ewaldType="trad"', 'ewaldType="trad" ewaldImages="3"'), 2)
    assert "ewaldNDim" in _refused(FERMIONS.replace('ewaldType="trad"', 'ewaldType="trad" ewaldNDim="1"'), 2)
