"""ctypes view of the C ABI in include/piqmc_hip.h and the workload builders.

The structures mirror the header field for field.  ``load_library`` opens the
in-tree HIP library (pi-qmc_amd/libpiqmc_hip.so) and raises if it is missing:
there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpiqmc_hip.so")

PQ_MAX_SPECIES = 4
PQ_MAX_SHO = 4
PQ_MAX_NODES = 4
PQ_MAX_SAMPLES = 8
PQ_MAX_MOVING = 8
PQ_MAX_NODE_PART = 32
PQ_MOVER_FREE = 0
PQ_MOVER_FREE_PBC = 1
PQ_NODES_FREE = 0
PQ_NODES_SHO = 1
PQ_OK = 0
PQ_ERR_ARG = -1
PQ_ERR_HIP = -2
PQ_ERR_NOMEM = -3
PQ_ERR_STATE = -4
PQ_CHOOSER_SPECIES = 0
PQ_CHOOSER_WALKING = 1
PQ_CHOOSER_ALL = 2
PQ_CHOOSER_ALL_RANDOM = 3
PQ_PREC_FP64 = 0
PQ_PREC_MIXED = 1


class SpeciesDesc(C.Structure):
    _fields_ = [("count", C.c_int), ("mass", C.c_double), ("charge", C.c_double),
                ("is_static", C.c_int)]


class ShoDesc(C.Structure):
    _fields_ = [("species", C.c_int), ("omega", C.c_double),
                ("center", C.c_double * 3), ("ndim", C.c_int)]


class CoulombDesc(C.Structure):
    _fields_ = [("enabled", C.c_int), ("epsilon", C.c_double), ("norder", C.c_int),
                ("rmin", C.c_double), ("rmax", C.c_double), ("ngpts", C.c_int),
                ("use_ewald", C.c_int), ("ewald_rcut", C.c_double), ("ewald_kcut", C.c_double),
                ("ewald_type", C.c_int), ("ewald_kappa", C.c_double)]


class NodeDesc(C.Structure):
    _fields_ = [("species", C.c_int), ("model", C.c_int), ("omega", C.c_double),
                ("temperature", C.c_double), ("nodal_action", C.c_int),
                ("many_body_distance", C.c_int), ("use_updates", C.c_int)]


class SampleDesc(C.Structure):
    _fields_ = [("species", C.c_int), ("nmoving", C.c_int), ("chooser", C.c_int),
                ("nrepeat", C.c_int), ("delayed_rejection", C.c_int),
                ("sigma_factor", C.c_double), ("new_sigma_factor", C.c_double)]


def _fill_sample(dst, smp):
    """smp: (species, npart[, chooser[, nrepeat[, opts]]]) with opts a dict of
    delayed_rejection, sigma_factor (defaultSigmaFactor) and new_sigma_factor."""
    dst.species, dst.nmoving = smp[0], smp[1]
    dst.chooser = smp[2] if len(smp) > 2 else PQ_CHOOSER_SPECIES
    dst.nrepeat = smp[3] if len(smp) > 3 else 1
    opts = smp[4] if len(smp) > 4 else {}
    dst.delayed_rejection = 1 if opts.get("delayed_rejection") else 0
    dst.sigma_factor = opts.get("sigma_factor", 0.0)
    dst.new_sigma_factor = opts.get("new_sigma_factor", 0.0)


PQ_MAX_GROUPS = 4


class SectionGroup(C.Structure):
    _fields_ = [("nlevel", C.c_int), ("nsample", C.c_int),
                ("sample", SampleDesc * PQ_MAX_SAMPLES), ("nsection", C.c_int)]


class ChainDesc(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nslice", C.c_int), ("nlevel", C.c_int),
                ("nspecies", C.c_int), ("tau", C.c_double), ("cell", C.c_double * 3),
                ("species", SpeciesDesc * PQ_MAX_SPECIES), ("spring", C.c_int),
                ("nsho", C.c_int), ("sho", ShoDesc * PQ_MAX_SHO),
                ("coulomb", CoulombDesc), ("nnodes", C.c_int),
                ("nodes", NodeDesc * PQ_MAX_NODES), ("mover", C.c_int),
                ("nsample", C.c_int), ("sample", SampleDesc * PQ_MAX_SAMPLES),
                ("nsection", C.c_int), ("ngroup", C.c_int),
                ("group", SectionGroup * (PQ_MAX_GROUPS - 1))]

    @property
    def npart(self):
        return sum(self.species[s].count for s in range(self.nspecies))

    def nbeads(self):
        return self.nslice * self.npart * self.ndim


PQ_XML_MAX_CF_SPECIES = 8


class XmlInput(C.Structure):
    """pq_xml_input: a pi-qmc input file read for the device path."""
    _fields_ = [("sweep", ChainDesc), ("nmeasure", C.c_int), ("nblock", C.c_int),
                ("has_thermalize", C.c_int), ("thermalize", ChainDesc), ("nthermalize", C.c_int),
                ("has_seed", C.c_int), ("iseed", C.c_int), ("read_paths", C.c_char * 256),
                ("read_bfactor", C.c_int), ("thermal_energy", C.c_int), ("virial_nwindow", C.c_int),
                ("coulomb_energy", C.c_int), ("coulomb_epsilon", C.c_double), ("coulomb_ewald", C.c_int),
                ("coulomb_rcut", C.c_double), ("coulomb_kcut", C.c_double), ("coulomb_ewald_type", C.c_int),
                ("coulomb_kappa", C.c_double), ("pair_cf_nspecies", C.c_int),
                ("pair_cf_species", C.c_int * PQ_XML_MAX_CF_SPECIES), ("pair_cf_rmin", C.c_double),
                ("pair_cf_rmax", C.c_double), ("pair_cf_nbin", C.c_int), ("nignored", C.c_int)]


def read_xml(path_or_text, ndim=3, text=False):
    """pq_read_xml (a file) or pq_parse_xml (text=True): an XmlInput; raises
    PQError with the reader's message on inputs the device path does not model."""
    lib = load_library()
    out = XmlInput()
    fn = lib.pq_parse_xml if text else lib.pq_read_xml
    check(fn(path_or_text.encode() if isinstance(path_or_text, str) else path_or_text, ndim, C.byref(out)))
    return out


class ChainStats(C.Structure):
    _fields_ = [("draws", C.c_uint64), ("bead_updates", C.c_uint64),
                ("accepted_bead_updates", C.c_uint64), ("tries", C.c_uint64 * 8),
                ("accepts", C.c_uint64 * 8), ("node_errors", C.c_uint64),
                ("sweeps", C.c_uint64), ("t_start", C.c_uint64), ("t_end", C.c_uint64)]


def make_desc(ndim, nslice, nlevel, tau, species, cell=1000.0, spring=True,
              sho=(), coulomb=None, nodes=(), mover=PQ_MOVER_FREE, samples=None,
              nsection=None):
    """species: [(count, mass, charge[, is_static])]; sho: [(species, omega)];
    coulomb: dict(epsilon, norder, rmin, rmax, ngpts[, ewald, ewald_rcut,
    ewald_kcut]) or None;
    nodes: [(species, model, omega[, use_updates])]; samples [(species, npart[, chooser[, nrepeat[,
    opts]]])] with opts {delayed_rejection, sigma_factor, new_sigma_factor},
    default one npart=1 SpeciesParticleChooser per species;
    nsection default M / 2^L (one sweep touches every slice once on average)."""
    d = ChainDesc()
    d.ndim, d.nslice, d.nlevel, d.nspecies, d.tau = ndim, nslice, nlevel, len(species), tau
    cells = cell if isinstance(cell, (list, tuple)) else [cell] * 3
    for k in range(3):
        d.cell[k] = cells[k]
    for s, sp in enumerate(species):
        d.species[s].count, d.species[s].mass, d.species[s].charge = sp[0], sp[1], sp[2]
        d.species[s].is_static = 1 if len(sp) > 3 and sp[3] else 0
    d.spring = 1 if spring else 0
    d.nsho = len(sho)
    for k, (s, w) in enumerate(sho):
        d.sho[k].species, d.sho[k].omega, d.sho[k].ndim = s, w, 0
    if coulomb:
        d.coulomb.enabled = 1
        d.coulomb.epsilon = coulomb.get("epsilon", 1.0)
        d.coulomb.norder = coulomb.get("norder", 3)
        d.coulomb.rmin = coulomb.get("rmin", 0.0)
        d.coulomb.rmax = coulomb.get("rmax", 0.0)
        d.coulomb.ngpts = coulomb.get("ngpts", 0)
        d.coulomb.use_ewald = 1 if coulomb.get("ewald") else 0
        d.coulomb.ewald_rcut = coulomb.get("ewald_rcut", 0.0)
        d.coulomb.ewald_kcut = coulomb.get("ewald_kcut", 0.0)
        d.coulomb.ewald_type = 1 if coulomb.get("ewald_type", "opt") == "trad" else 0
        d.coulomb.ewald_kappa = coulomb.get("ewald_kappa", 0.0)
    d.nnodes = len(nodes)
    for k, nd in enumerate(nodes):
        s, model, w = nd[:3]
        d.nodes[k].species, d.nodes[k].model, d.nodes[k].omega = s, model, w
        d.nodes[k].temperature, d.nodes[k].nodal_action = 0.0, 1
        d.nodes[k].use_updates = 1 if len(nd) > 3 and nd[3] else 0
    d.mover = mover
    if samples is None:
        samples = [(s, 1) for s in range(len(species))]
    d.nsample = len(samples)
    for k, smp in enumerate(samples):
        _fill_sample(d.sample[k], smp)
    d.nsection = nsection if nsection is not None else max(1, nslice >> nlevel)
    return d


def add_group(d, nlevel, samples, nsection):
    """Append a <ChooseSection nlevel> group (samples as for make_desc,
    repeated nsection times) to a sweep."""
    g = d.group[d.ngroup]
    g.nlevel, g.nsample, g.nsection = nlevel, len(samples), nsection
    for k, smp in enumerate(samples):
        _fill_sample(g.sample[k], smp)
    d.ngroup += 1
    return d


# ----------------------------------------------------------- workloads

def dot3d_2e_desc():
    """2 e- (up, dn) in a 3D harmonic dot, omega=1, Coulomb 1/r, beta=10, M=200, L=3."""
    return make_desc(3, 200, 3, 0.05, [(1, 1.0, -1.0), (1, 1.0, -1.0)],
                     sho=[(0, 1.0), (1, 1.0)],
                     coulomb=dict(epsilon=1.0, norder=3, rmin=0.005, rmax=50.0, ngpts=1024))


def dot2d_3p3_desc():
    """3+3 e- in a 2D dot, omega=1, Coulomb eps=1, SHONodes, beta=8, M=400, L=4."""
    return make_desc(2, 400, 4, 0.02, [(3, 1.0, -1.0), (3, 1.0, -1.0)],
                     sho=[(0, 1.0), (1, 1.0)],
                     coulomb=dict(epsilon=1.0, norder=3, rmin=0.005, rmax=50.0, ngpts=1024),
                     nodes=[(0, PQ_NODES_SHO, 1.0), (1, PQ_NODES_SHO, 1.0)])


def batch2d_descs(nchain=1024, seed=0x5EED):
    """Uneven batch of 2D dots: N~U{2..20}, M in {128..1024}, omega~U[0.5,2],
    SHONodes on even chains, FreeParticleNodes (box 20) on odd chains."""
    out = []
    for c in range(nchain):
        rng = np.random.default_rng(seed + c)
        n = int(rng.integers(2, 21))
        up, dn = (n + 1) // 2, n // 2
        m = int(rng.choice([128, 256, 512, 1024]))
        w = float(rng.uniform(0.5, 2.0))
        sp = [(up, 1.0, -1.0)] + ([(dn, 1.0, -1.0)] if dn else [])
        model = PQ_NODES_SHO if c % 2 == 0 else PQ_NODES_FREE
        nodes = [(s, model, w) for s in range(len(sp))]
        out.append(make_desc(2, m, 4, 0.02, sp, cell=20.0 if model == PQ_NODES_FREE else 1000.0,
                             sho=[(s, w) for s in range(len(sp))],
                             coulomb=dict(epsilon=1.0, norder=3, rmin=0.005, rmax=50.0, ngpts=1024),
                             nodes=nodes))
    return out


def ehbox_desc():
    """32 e (m=1) + 32 h (m=2), spin-polarised, cubic box L=10, min-image Coulomb,
    FreeParticleNodes 32x32 per species, beta=20, M=1024, L=5, FreePBC mover."""
    return make_desc(3, 1024, 5, 20.0 / 1024, [(32, 1.0, -1.0), (32, 2.0, 1.0)], cell=10.0,
                     coulomb=dict(epsilon=1.0, norder=3, rmin=0.0, rmax=float(np.sqrt(3) * 10),
                                  ngpts=1024),
                     nodes=[(0, PQ_NODES_FREE, 0.0), (1, PQ_NODES_FREE, 0.0)],
                     mover=PQ_MOVER_FREE_PBC)


def batch3d_descs(nchain=4096, seed=0xC0FFEE):
    """Large uneven batch of 3D dots: N~U{2..64} in two species, M=16*U{16..128}, SHONodes."""
    out = []
    for c in range(nchain):
        rng = np.random.default_rng(seed + c)
        n = int(rng.integers(2, 65))
        up, dn = (n + 1) // 2, n // 2
        m = 16 * int(rng.integers(16, 129))
        sp = [(up, 1.0, -1.0), (dn, 1.0, -1.0)]
        out.append(make_desc(3, m, 4, 0.05, sp, sho=[(0, 1.0), (1, 1.0)],
                             coulomb=dict(epsilon=1.0, norder=3, rmin=0.005, rmax=50.0, ngpts=1024),
                             nodes=[(0, PQ_NODES_SHO, 1.0), (1, PQ_NODES_SHO, 1.0)]))
    return out


def initial_beads(desc, seed, replicate=None):
    """Seeded N(0,1) beads; replicated across slices when the chain has node
    actions (R(t) = R(t+beta/2) makes every determinant positive)."""
    rng = np.random.default_rng(seed)
    n, d, m = desc.npart, desc.ndim, desc.nslice
    if replicate is None:
        replicate = desc.nnodes > 0
    if replicate:
        if desc.mover == PQ_MOVER_FREE_PBC and desc.cell[0] < 100:
            r = rng.uniform(-0.5, 0.5, size=(n, d)) * np.array(desc.cell[:d])
        else:
            r = rng.standard_normal((n, d))
        b = np.broadcast_to(r, (m, n, d)).copy()
    else:
        b = rng.standard_normal((m, n, d))
    return np.ascontiguousarray(b, dtype=np.float64)


def descs_array(descs):
    arr = (ChainDesc * len(descs))()
    for k, d in enumerate(descs):
        arr[k] = d
    return arr


# ----------------------------------------------------------- library

_lib = None


def load_library(path=LIB_PATH):
    """Open the HIP library; raises OSError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    P = C.POINTER
    dp, ip = P(C.c_double), P(C.c_int)
    sig = {
        "pq_last_error": (C.c_char_p, []),
        "pq_batch_create": (C.c_int, [P(ChainDesc), C.c_int, C.c_int, C.c_int, P(C.c_void_p)]),
        "pq_batch_destroy": (None, [C.c_void_p]),
        "pq_batch_info": (C.c_int, [C.c_void_p, C.c_int, ip, ip, ip, P(C.c_int64)]),
        "pq_set_paths": (C.c_int, [C.c_void_p, C.c_int, dp]),
        "pq_get_paths": (C.c_int, [C.c_void_p, C.c_int, dp]),
        "pq_set_all_paths": (C.c_int, [C.c_void_p, dp, C.c_int64]),
        "pq_get_all_paths": (C.c_int, [C.c_void_p, dp, C.c_int64]),
        "pq_seed": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32]),
        "pq_set_uniforms": (C.c_int, [C.c_void_p, dp, C.c_int64]),
        "pq_run_sweeps": (C.c_int, [C.c_void_p, C.c_int]),
        "pq_sync": (C.c_int, [C.c_void_p]),
        "pq_set_schedule": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "pq_get_schedule": (C.c_int, [C.c_void_p, ip, ip, ip]),
        "pq_set_staging": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "pq_get_staging": (C.c_int, [C.c_void_p, ip, ip]),
        "pq_autotune": (C.c_int, [C.c_void_p, C.c_int, P(C.c_float), ip]),
        "pq_set_split": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "pq_get_split": (C.c_int, [C.c_void_p, ip, ip]),
        "pq_write_paths_file": (C.c_int, [C.c_char_p, dp, ip, C.c_int, C.c_int, C.c_int]),
        "pq_run_measured": (C.c_int, [C.c_void_p, C.c_int, C.c_int, dp]),
        "pq_read_paths_file": (C.c_int, [C.c_char_p, dp, ip, C.c_int, C.c_int, C.c_int, C.c_int]),
        "pq_set_permutation": (C.c_int, [C.c_void_p, C.c_int, ip]),
        "pq_write_ascii_report": (C.c_int, [C.c_char_p, C.c_int, P(C.c_char_p), P(C.c_char_p), dp, dp,
                                            C.c_int, dp]),
        "pq_get_permutation": (C.c_int, [C.c_void_p, C.c_int, ip]),
        "pq_write_all_paths": (C.c_int, [C.c_void_p, C.c_char_p]),
        "pq_read_all_paths": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
        "pq_run_sweeps_timed": (C.c_int, [C.c_void_p, C.c_int, P(C.c_float)]),
        "pq_get_all_stats": (C.c_int, [C.c_void_p, P(ChainStats)]),
        "pq_get_pair_table": (C.c_int, [C.c_void_p, C.c_int, C.c_int, ip, dp, dp, dp]),
        "pq_get_stats": (C.c_int, [C.c_void_p, C.c_int, P(ChainStats)]),
        "pq_reset_stats": (C.c_int, [C.c_void_p]),
        "pq_get_node_scale": (C.c_int, [C.c_void_p, C.c_int, dp]),
        "pq_measure_energy": (C.c_int, [C.c_void_p, dp]),
        "pq_measure_coulomb_energy": (C.c_int, [C.c_void_p, C.c_double, dp]),
        "pq_measure_virial_energy": (C.c_int, [C.c_void_p, C.c_int, dp]),
        "pq_measure_ewald_coulomb_energy": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, dp]),
        "pq_measure_trad_ewald_coulomb_energy": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                                           C.c_double, dp]),
        "pq_pair_histogram": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_int, P(C.c_uint64)]),
        "pq_pair_cf_histogram": (C.c_int, [C.c_void_p, C.c_int, ip, C.c_double, C.c_double,
                                           C.c_int, P(C.c_uint64)]),
        "pq_set_measure_pair_cf": (C.c_int, [C.c_void_p, C.c_int, ip, C.c_double, C.c_double, C.c_int]),
        "pq_collect_pair_cf": (C.c_int, [C.c_void_p, P(C.c_uint64)]),
        "pq_run_measured_timed": (C.c_int, [C.c_void_p, C.c_int, C.c_int, dp, P(C.c_float)]),
        "pq_set_sweep_plan": (C.c_int, [C.c_void_p, ip]),
        "pq_read_xml": (C.c_int, [C.c_char_p, C.c_int, P(XmlInput)]),
        "pq_parse_xml": (C.c_int, [C.c_char_p, C.c_int, P(XmlInput)]),
        "pq_get_sweep_plan": (C.c_int, [C.c_void_p, ip]),
        "pq_clock_hz": (C.c_int, [C.c_void_p, dp]),
        "pq_action_difference": (C.c_int, [C.c_void_p, C.c_int, ip, ip, dp, dp, ip, C.c_int, dp]),
        "pq_bead_utau": (C.c_int, [C.c_void_p, C.c_int, ip, ip, ip, dp]),
        "pq_bead_action": (C.c_int, [C.c_void_p, C.c_int, dp, ip, C.c_int, ip, ip, dp]),
        "pq_section_begin": (C.c_int, [C.c_void_p, C.c_int, dp, dp]),
        "pq_double_action_difference": (C.c_int, [C.c_void_p, C.c_int, C.c_int, dp, ip, C.c_int, dp]),
        "pq_accept_last_move": (C.c_int, [C.c_void_p, C.c_int]),
        "pq_node_evaluate": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, dp, dp, dp, ip, dp, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def write_ascii_report(filename, names, values, units=None, shift=None, scale=None):
    """pimc.dat of the reference's ASCII report: values [nblock][nscalar]."""
    lib = load_library()
    n = len(names)
    v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1, n)
    cn = (C.c_char_p * n)(*[x.encode() for x in names])
    cu = (C.c_char_p * n)(*[(u or "").encode() for u in units]) if units else None
    sh = np.ascontiguousarray(shift, dtype=np.float64) if shift is not None else None
    sc = np.ascontiguousarray(scale, dtype=np.float64) if scale is not None else None
    check(lib.pq_write_ascii_report(filename.encode(), n, cn, cu, dptr(sh) if sh is not None else None,
                                    dptr(sc) if sc is not None else None, v.shape[0], dptr(v)))


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class PQError(RuntimeError):
    pass


def check(rc, lib=None):
    if rc != 0:
        lib = lib or load_library()
        msg = lib.pq_last_error()
        raise PQError(f"pq call failed ({rc}): {msg.decode() if msg else ''}")


class Batch:
    """Owning handle of a pq_batch (device state of many chains)."""

    def __init__(self, descs, precision=PQ_PREC_FP64, device=0):
        self.lib = load_library()
        self.descs = list(descs)
        self._arr = descs_array(self.descs)
        h = C.c_void_p()
        check(self.lib.pq_batch_create(self._arr, len(self.descs), precision, device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.pq_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_paths(self, c, beads):
        b = np.ascontiguousarray(beads, dtype=np.float64)
        check(self.lib.pq_set_paths(self.h, c, dptr(b)))

    def get_paths(self, c):
        d = self.descs[c]
        b = np.empty((d.nslice, d.npart, d.ndim))
        check(self.lib.pq_get_paths(self.h, c, dptr(b)))
        return b

    def set_permutation(self, c, perm):
        p = np.ascontiguousarray(perm, dtype=np.int32)
        check(self.lib.pq_set_permutation(self.h, c, iptr(p)))

    def get_permutation(self, c):
        p = np.zeros(self.descs[c].npart, dtype=np.int32)
        check(self.lib.pq_get_permutation(self.h, c, iptr(p)))
        return p

    def seed(self, seed, first_stream=0):
        check(self.lib.pq_seed(self.h, seed, first_stream))

    def set_uniforms(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        check(self.lib.pq_set_uniforms(self.h, dptr(u), u.shape[-1]))

    def run_sweeps(self, n):
        check(self.lib.pq_run_sweeps(self.h, n))

    def sync(self):
        check(self.lib.pq_sync(self.h))

    def set_schedule(self, team=-1, inv_lds=-1):
        check(self.lib.pq_set_schedule(self.h, team, inv_lds))

    def schedule(self):
        t, i, b = C.c_int(), C.c_int(), C.c_int()
        check(self.lib.pq_get_schedule(self.h, C.byref(t), C.byref(i), C.byref(b)))
        return t.value, i.value, b.value

    def set_staging(self, tables_lds=-1, state_lds=-1):
        check(self.lib.pq_set_staging(self.h, tables_lds, state_lds))

    def staging(self):
        t, f = C.c_int(), C.c_int()
        check(self.lib.pq_get_staging(self.h, C.byref(t), C.byref(f)))
        return t.value, f.value

    def set_split(self, heavy_chains=0, heavy_team=256):
        check(self.lib.pq_set_split(self.h, heavy_chains, heavy_team))

    def split(self):
        n, t = C.c_int(), C.c_int()
        check(self.lib.pq_get_split(self.h, C.byref(n), C.byref(t)))
        return n.value, t.value

    def autotune(self, nsweep=1):
        """Time every distinct schedule over nsweep sweeps and keep the fastest;
        returns (ms per sweep of the fastest, sweeps run).  The sweeps advance
        the chains."""
        ms, n = C.c_float(), C.c_int()
        check(self.lib.pq_autotune(self.h, nsweep, C.byref(ms), C.byref(n)))
        return ms.value, n.value * nsweep

    def bead_action(self, c, ipart, islice, beads=None, perm=None):
        """getBeadAction of beads (ipart[q], islice[q]) of chain c: [nq][3 + 2D]
        = u, utau, ulambda, fm, fp; on the caller's paths (beads [M][N][D],
        perm) or, with beads None, on the device-resident paths."""
        ii = np.ascontiguousarray(ipart, dtype=np.int32)
        tt = np.ascontiguousarray(islice, dtype=np.int32)
        D = self.descs[c].ndim
        out = np.zeros((ii.size, 3 + 2 * D))
        bp = dptr(np.ascontiguousarray(beads, dtype=np.float64)) if beads is not None else None
        keep = None
        if perm is not None:
            keep = np.ascontiguousarray(perm, dtype=np.int32)
        if beads is not None:
            self._bead_keep = np.ascontiguousarray(beads, dtype=np.float64)
            bp = dptr(self._bead_keep)
        check(self.lib.pq_bead_action(self.h, c, bp, iptr(keep) if keep is not None else None, ii.size,
                                      iptr(ii), iptr(tt), dptr(out)))
        return out

    def set_sweep_plan(self, sweeps=None):
        """Sweeps per launch of each chain (None: one each)."""
        if sweeps is None:
            check(self.lib.pq_set_sweep_plan(self.h, None))
            return
        p = np.ascontiguousarray(sweeps, dtype=np.int32)
        check(self.lib.pq_set_sweep_plan(self.h, iptr(p)))

    def sweep_plan(self):
        p = np.zeros(len(self.descs), dtype=np.int32)
        check(self.lib.pq_get_sweep_plan(self.h, iptr(p)))
        return p

    def clock_hz(self):
        v = C.c_double()
        check(self.lib.pq_clock_hz(self.h, C.byref(v)))
        return v.value

    def all_stats(self):
        st = (ChainStats * len(self.descs))()
        check(self.lib.pq_get_all_stats(self.h, st))
        return st

    def stats(self, c):
        st = ChainStats()
        check(self.lib.pq_get_stats(self.h, c, C.byref(st)))
        return st

    def energy(self):
        e = np.empty(len(self.descs))
        check(self.lib.pq_measure_energy(self.h, dptr(e)))
        return e

    def coulomb_energy(self, epsilon=1.0):
        e = np.empty(len(self.descs))
        check(self.lib.pq_measure_coulomb_energy(self.h, epsilon, dptr(e)))
        return e

    def run_measured(self, nblock, nsweep):
        """nblock x (nsweep sweeps, thermodynamic energy) in one captured graph:
        energies [nblock][nchain]."""
        e = np.empty((nblock, len(self.descs)))
        check(self.lib.pq_run_measured(self.h, nblock, nsweep, dptr(e)))
        return e

    def pair_histogram(self, s1, s2, rmin, rmax, nbin):
        """Integer g(r) counts [nchain][nbin] of the current paths."""
        h = np.zeros((len(self.descs), nbin), dtype=np.uint64)
        check(self.lib.pq_pair_histogram(self.h, s1, s2, rmin, rmax, nbin,
                                         h.ctypes.data_as(C.POINTER(C.c_uint64))))
        return h

    def pair_cf_histogram(self, species, rmin, rmax, nbin):
        """species[0] against the partner species species[1:]: counts [nchain][nbin]."""
        sp = np.ascontiguousarray(species, dtype=np.int32)
        h = np.zeros((len(self.descs), nbin), dtype=np.uint64)
        check(self.lib.pq_pair_cf_histogram(self.h, len(sp), iptr(sp), rmin, rmax, nbin,
                                            h.ctypes.data_as(C.POINTER(C.c_uint64))))
        return h

    def set_measure_pair_cf(self, species, rmin=0.0, rmax=1.0, nbin=1):
        """Pair-correlation estimator of every run_measured Measure (species [] removes it)."""
        sp = np.ascontiguousarray(species if len(species) else [0], dtype=np.int32)
        check(self.lib.pq_set_measure_pair_cf(self.h, len(species), iptr(sp), rmin, rmax, nbin))
        self._cf_nbin = nbin if len(species) else 0

    def collect_pair_cf(self):
        h = np.zeros((len(self.descs), self._cf_nbin), dtype=np.uint64)
        check(self.lib.pq_collect_pair_cf(self.h, h.ctypes.data_as(C.POINTER(C.c_uint64))))
        return h

    def run_measured_timed(self, nblock, nsweep):
        """run_measured with each block's sampler-launch time (ms) from events in the graph."""
        e = np.empty((nblock, len(self.descs)))
        ms = (C.c_float * nblock)()
        check(self.lib.pq_run_measured_timed(self.h, nblock, nsweep, dptr(e), ms))
        return e, np.array(ms[:], dtype=np.float64)

    def ewald_coulomb_energy(self, epsilon, rcut, kcut, trad=False, kappa=0.0):
        e = np.empty(len(self.descs))
        if trad:
            check(self.lib.pq_measure_trad_ewald_coulomb_energy(self.h, epsilon, rcut, kcut, kappa, dptr(e)))
        else:
            check(self.lib.pq_measure_ewald_coulomb_energy(self.h, epsilon, rcut, kcut, dptr(e)))
        return e

    def virial_energy(self, nwindow=1):
        e = np.empty(len(self.descs))
        check(self.lib.pq_measure_virial_energy(self.h, nwindow, dptr(e)))
        return e
