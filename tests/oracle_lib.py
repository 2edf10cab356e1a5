"""ctypes binding of the oracle (oracle/libpq_oracle.so), the CPU checker.

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "pi-qmc_amd"))
import piqmc_capi as capi  # noqa: E402

ORACLE_PATH = os.path.join(ROOT, "oracle", "libpq_oracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libpq_ref.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_PATH):
        raise OSError(f"oracle not built: {ORACLE_PATH}")
    lib = C.CDLL(ORACLE_PATH)
    P = C.POINTER
    dp, ip, vp = P(C.c_double), P(C.c_int), C.c_void_p
    sig = {
        "oq_glibc_sequence": (None, [C.c_uint, C.c_int, P(C.c_int32)]),
        "oq_philox_raw": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "oq_philox_u32": (C.c_uint32, [C.c_uint64, C.c_uint32, C.c_uint64]),
        "oq_philox_uniform": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint64]),
        "oq_coulomb1d": (C.c_double, [C.c_double, C.c_int, C.c_double]),
        "oq_coulomb3d": (C.c_double, [C.c_double, C.c_int, C.c_double]),
        "oq_pg_nterms": (C.c_int, [C.c_double, C.c_double]),
        "oq_pg_eval": (None, [C.c_double, C.c_double, C.c_double, dp]),
        "oq_pbc": (None, [dp, C.c_int, dp]),
        "oq_create": (vp, [P(capi.ChainDesc)]),
        "oq_destroy": (None, [vp]),
        "oq_set_beads": (None, [vp, dp]),
        "oq_get_beads": (None, [vp, dp]),
        "oq_set_permutation": (C.c_int, [vp, ip]),
        "oq_get_permutation": (None, [vp, ip]),
        "oq_perm_compose": (None, [ip, ip, C.c_int, C.c_int]),
        "oq_npair": (C.c_int, [vp]),
        "oq_pair_table": (C.c_int, [vp, C.c_int, ip, dp, dp, dp]),
        "oq_rng_glibc": (None, [vp, C.c_uint]),
        "oq_rng_philox": (None, [vp, C.c_uint64, C.c_uint32]),
        "oq_rng_inject": (None, [vp, dp, C.c_int64]),
        "oq_rng_next": (C.c_double, [vp]),
        "oq_rng_count": (C.c_uint64, [vp]),
        "oq_rng_exhausted": (C.c_int, [vp]),
        "oq_action_difference": (C.c_double, [vp, dp, dp, ip, C.c_int, C.c_int, C.c_int]),
        "oq_bead_utau": (C.c_double, [vp, C.c_int, C.c_int, C.c_int]),
        "oq_node_evaluate": (C.c_int, [vp, C.c_int, dp, dp, C.c_int, C.c_int, dp]),
        "oq_node_distance": (None, [vp, C.c_int, dp, dp, C.c_int, dp, dp]),
        "oq_node_scale": (C.c_double, [vp, C.c_int]),
        "oq_fn_initialize": (None, [vp, dp, dp]),
        "oq_fn_action_difference": (C.c_double, [vp, dp, ip, C.c_int, C.c_int]),
        "oq_fn_accept": (None, [vp]),
        "oq_fn_state": (None, [vp, C.c_int, C.c_int, dp, dp]),
        "oq_mover": (C.c_double, [vp, dp, dp, ip, C.c_int, C.c_int]),
        "oq_run_sweeps": (None, [vp, C.c_int]),
        "oq_get_stats": (None, [vp, P(capi.ChainStats)]),
        "oq_thermo_energy": (C.c_double, [vp]),
        "oq_coulomb_energy": (C.c_double, [vp, C.c_double]),
        "oq_ewald_coulomb_energy": (C.c_double, [vp, C.c_double, C.c_double, C.c_double]),
        "oq_ewald_coulomb_energy_type": (C.c_double, [vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double]),
        "oq_trad_estimator_kappa": (C.c_double, [vp]),
        "oq_ewald_info": (C.c_int, [vp, dp, ip, dp, C.c_int, ip, dp]),
        "oq_ewald_long_range": (C.c_double, [vp, dp]),
        "oq_ewald_fr": (C.c_double, [vp, C.c_double]),
        "oq_ref_erf": (C.c_double, [C.c_double]),
        "oq_bead_action": (None, [vp, C.c_int, C.c_int, dp, dp, dp, dp]),
        "oq_virial_energy": (C.c_double, [vp, C.c_int]),
        "oq_energy_blocks": (None, [vp, C.c_int, C.c_int, dp]),
        "oq_pair_histogram": (None, [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                     P(C.c_uint64)]),
        "oq_pair_cf_histogram": (None, [vp, C.c_int, ip, C.c_double, C.c_double, C.c_int,
                                        P(C.c_uint64)]),
        "oq_run_batch": (C.c_uint64, [P(capi.ChainDesc), C.c_int, dp, C.c_int, C.c_uint64,
                                      C.c_uint32, C.c_int, dp]),
        "oq_run_batch_measured": (C.c_uint64, [P(capi.ChainDesc), C.c_int, dp, C.c_int, C.c_uint64,
                                               C.c_uint32, C.c_int, C.c_int, ip, C.c_double, C.c_double,
                                               C.c_int, dp, P(C.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def load_ref():
    """The reference's own Coulomb/PeriodicGaussian/erf code built into oracle/_ref."""
    if not os.path.exists(REF_PATH):
        return None
    lib = C.CDLL(REF_PATH)
    for name in ("ref_coulomb1d", "ref_coulomb3d"):
        f = getattr(lib, name)
        f.restype = C.c_double
        f.argtypes = [C.c_double, C.c_int, C.c_double]
    lib.ref_pg_eval.restype = None
    lib.ref_pg_eval.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
    lib.ref_pg_nterms.restype = C.c_int
    lib.ref_pg_nterms.argtypes = [C.c_double, C.c_double]
    lib.ref_erf.restype = C.c_double
    lib.ref_erf.argtypes = [C.c_double]
    return lib


dptr = capi.dptr
iptr = capi.iptr


class Chain:
    """One oracle chain (owning)."""

    def __init__(self, desc):
        self.lib = load()
        self.desc = desc
        self._d = capi.ChainDesc()
        C.pointer(self._d)[0] = desc
        self.h = self.lib.oq_create(C.byref(self._d))
        if not self.h:
            raise ValueError("configuration not supported by the oracle")
        self.N, self.D, self.M = desc.npart, desc.ndim, desc.nslice
        self.S = (1 << desc.nlevel) + 1
        self._keep = None

    def __del__(self):
        try:
            self.lib.oq_destroy(self.h)
        except Exception:
            pass

    def set_beads(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        self.lib.oq_set_beads(self.h, dptr(b))

    def get_beads(self):
        b = np.empty((self.M, self.N, self.D))
        self.lib.oq_get_beads(self.h, dptr(b))
        return b

    def set_permutation(self, p):
        p = np.ascontiguousarray(p, dtype=np.int32)
        if self.lib.oq_set_permutation(self.h, iptr(p)) != 0:
            raise ValueError("not a permutation")

    def get_permutation(self):
        p = np.zeros(self.N, dtype=np.int32)
        self.lib.oq_get_permutation(self.h, iptr(p))
        return p

    def rng_glibc(self, seed):
        self.lib.oq_rng_glibc(self.h, seed)

    def rng_philox(self, seed, stream):
        self.lib.oq_rng_philox(self.h, seed, stream)

    def rng_inject(self, u):
        self._keep = np.ascontiguousarray(u, dtype=np.float64)
        self.lib.oq_rng_inject(self.h, dptr(self._keep), self._keep.size)

    def action_difference(self, sec, mov, idx, level, mask=7):
        sec = np.ascontiguousarray(sec, dtype=np.float64)
        mov = np.ascontiguousarray(mov, dtype=np.float64)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        return self.lib.oq_action_difference(self.h, dptr(sec), dptr(mov), iptr(idx),
                                              len(idx), level, mask)

    def bead_utau(self, i, t, mask=7):
        return self.lib.oq_bead_utau(self.h, i, t, mask)

    def _check_stream(self):
        if self.lib.oq_rng_exhausted(self.h):
            raise RuntimeError(f"injected uniform stream exhausted after {self.lib.oq_rng_count(self.h)} draws")

    def run_sweeps(self, n):
        self.lib.oq_run_sweeps(self.h, n)
        self._check_stream()

    def stats(self):
        st = capi.ChainStats()
        self.lib.oq_get_stats(self.h, C.byref(st))
        return st

    def energy(self):
        return self.lib.oq_thermo_energy(self.h)

    def coulomb_energy(self, epsilon=1.0):
        return self.lib.oq_coulomb_energy(self.h, epsilon)

    def ewald_coulomb_energy(self, epsilon, rcut, kcut):
        return self.lib.oq_ewald_coulomb_energy(self.h, epsilon, rcut, kcut)

    def trad_ewald_coulomb_energy(self, epsilon, rcut, kcut, kappa=0.0):
        return self.lib.oq_ewald_coulomb_energy_type(self.h, epsilon, rcut, kcut, 1, kappa)

    def trad_estimator_kappa(self):
        return self.lib.oq_trad_estimator_kappa(self.h)

    def ewald_info(self):
        """(coef[8], nk, self energy, kv [nk][3], vk [nk]) of the CoulombAction's sum."""
        coef, nk, se = np.zeros(8), C.c_int(), C.c_double()
        if not self.lib.oq_ewald_info(self.h, dptr(coef), C.byref(nk), C.byref(se), -1, None, None):
            return None
        kv, vk = np.zeros((nk.value, 3), dtype=np.int32), np.zeros(nk.value)
        for k in range(nk.value):
            v, w = np.zeros(3, dtype=np.int32), C.c_double()
            self.lib.oq_ewald_info(self.h, None, None, None, k, iptr(v), C.byref(w))
            kv[k], vk[k] = v, w.value
        return coef, nk.value, se.value, kv, vk

    def ewald_long_range(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        return self.lib.oq_ewald_long_range(self.h, dptr(r))

    def ewald_fr(self, r):
        return self.lib.oq_ewald_fr(self.h, r)

    def bead_action(self, i, t):
        u, ut = C.c_double(), C.c_double()
        fm, fp = np.zeros(3), np.zeros(3)
        self.lib.oq_bead_action(self.h, i, t, C.byref(u), C.byref(ut), dptr(fm), dptr(fp))
        return u.value, ut.value, fm[:self.D], fp[:self.D]

    def virial_energy(self, nwindow=1):
        return self.lib.oq_virial_energy(self.h, nwindow)

    def energy_blocks(self, nblock, nmeasure):
        out = np.zeros(nblock)
        self.lib.oq_energy_blocks(self.h, nblock, nmeasure, dptr(out))
        self._check_stream()
        return out

    def pair_table(self, ip):
        meta = np.zeros(4, dtype=np.int32)
        grid = np.zeros(2)
        self.lib.oq_pair_table(self.h, ip, iptr(meta), dptr(grid), None, None)
        n = int(meta[3]) * (int(meta[2]) + 1)
        u, ut = np.zeros(n), np.zeros(n)
        self.lib.oq_pair_table(self.h, ip, iptr(meta), dptr(grid), dptr(u), dptr(ut))
        return meta, grid, u.reshape(meta[3], -1), ut.reshape(meta[3], -1)

    def fn_initialize(self, sec1, sec2):
        sec1 = np.ascontiguousarray(sec1, dtype=np.float64)
        sec2 = np.ascontiguousarray(sec2, dtype=np.float64)
        self.lib.oq_fn_initialize(self.h, dptr(sec1), dptr(sec2))

    def fn_action_difference(self, mov, idx, level):
        mov = np.ascontiguousarray(mov, dtype=np.float64)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        return self.lib.oq_fn_action_difference(self.h, dptr(mov), iptr(idx), len(idx), level)

    def fn_accept(self):
        self.lib.oq_fn_accept(self.h)

    def node_scale(self, inode):
        return self.lib.oq_node_scale(self.h, inode)

    def fn_state(self, inode, new=False):
        """(determinants [S], distances [S][2][ns]) of node action inode:
        dmValue/dist, or newDMValue/newDist with new=True"""
        S = (1 << self.desc.nlevel) + 1
        ns = self.desc.species[self.desc.nodes[inode].species].count
        det, dist = np.zeros(S), np.zeros(S * 2 * ns)
        self.lib.oq_fn_state(self.h, inode, 1 if new else 0, dptr(det), dptr(dist))
        return det, dist.reshape(S, 2, ns)

    def node_evaluate(self, inode, r1, r2, islice, scale_magnitude=0):
        r1 = np.ascontiguousarray(r1, dtype=np.float64)
        r2 = np.ascontiguousarray(r2, dtype=np.float64)
        det = C.c_double()
        err = self.lib.oq_node_evaluate(self.h, inode, dptr(r1), dptr(r2), islice,
                                        scale_magnitude, C.byref(det))
        return det.value, err

    def node_distance(self, inode, r1, r2, islice, ns):
        r1 = np.ascontiguousarray(r1, dtype=np.float64)
        r2 = np.ascontiguousarray(r2, dtype=np.float64)
        d1, d2 = np.zeros(ns), np.zeros(ns)
        self.lib.oq_node_distance(self.h, inode, dptr(r1), dptr(r2), islice, dptr(d1), dptr(d2))
        return d1, d2

    def mover(self, sec, mov, idx, level):
        sec = np.ascontiguousarray(sec, dtype=np.float64)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        assert mov.flags.c_contiguous and mov.dtype == np.float64
        v = self.lib.oq_mover(self.h, dptr(sec), dptr(mov), iptr(idx), len(idx), level)
        self._check_stream()
        return v

    def pair_histogram(self, s1, s2, rmin, rmax, nbin):
        h = np.zeros(nbin, dtype=np.uint64)
        self.lib.oq_pair_histogram(self.h, s1, s2, rmin, rmax, nbin,
                                   h.ctypes.data_as(C.POINTER(C.c_uint64)))
        return h

    def pair_cf_histogram(self, species, rmin, rmax, nbin):
        """species[0] against the partner list species[1:]."""
        sp = np.ascontiguousarray(species, dtype=np.int32)
        h = np.zeros(nbin, dtype=np.uint64)
        self.lib.oq_pair_cf_histogram(self.h, len(sp), iptr(sp), rmin, rmax, nbin,
                                      h.ctypes.data_as(C.POINTER(C.c_uint64)))
        return h
