"""Shared inputs for the GPU parity tests: seeded sections and moving beads,
and the glibc rand() uniform stream the reference would consume."""
import numpy as np

import oracle_lib as ol
import piqmc_capi as capi


def glibc_uniforms(seed, n):
    """rand()*(1/RAND_MAX) of glibc's TYPE_3 generator seeded with srand(seed)."""
    import ctypes as C
    out = (C.c_int32 * n)()
    ol.load().oq_glibc_sequence(seed, n, out)
    return np.frombuffer(out, dtype=np.int32).astype(np.float64) * (1.0 / 2147483647)


def random_section(desc, rng, spread=1.0, level_jitter=0.3):
    """A section of S slices (N(0,spread) per particle, smooth in time) and a
    moving copy of particle idx with every interior bead displaced."""
    S = (1 << desc.nlevel) + 1
    n, d = desc.npart, desc.ndim
    base = rng.standard_normal((1, n, d)) * spread
    sec = base + 0.3 * rng.standard_normal((S, n, d))
    idx = int(rng.integers(0, n))
    mov = sec[:, idx:idx + 1, :].copy()
    mov[1:-1] += level_jitter * rng.standard_normal((S - 2, 1, d))
    return np.ascontiguousarray(sec), np.ascontiguousarray(mov), idx


def rel_close(a, b, rtol, scale=1.0):
    """|a-b| <= rtol * max(|b|, scale): relative, with an absolute floor set by
    the magnitude of the terms that were summed."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), scale))


def small_dot2d_3p3(nslice=64):
    d = capi.dot2d_3p3_desc()
    d.nslice = nslice
    d.nsection = nslice >> d.nlevel
    return d
