"""Independent pins of oracle functions that the reference's known answers do
not reach:

* SHONodes / FreeParticleNodes determinants and nodal distances for 3..32
  particles against matrices built in numpy from the closed-form elements
  (SHONodes.cc:51-128,130-170, FreeParticleNodes.cc:72-157,184-234), with
  numpy.linalg.det and numpy.linalg.inv: the distances are sums over the
  inverse, so they pin it too;
* PairCFEstimator<1> with PairRadial (PairCFEstimator.h:72-93,
  PairDistance.h:22-39) against a numpy histogram of brute-force distances:
  periodic images, bin edges, r < min, r >= max, same-species pairs, several
  partner species;
* FreeMoverPBC (FreeMoverPBC.cc:59-162) by the statistics it must produce:
  one free particle on a 2D torus, whose winding numbers and thermal energy
  have closed forms.

None of these restates the oracle's arithmetic: the node matrices come from
the physics formulas and LAPACK-equivalent numpy routines, the histogram from
numpy.histogram on a nearest-image search, the torus from its partition
function."""
import itertools
import math

import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi


# ---------------------------------------------------------- node matrices

def _image_sum_gaussian(alpha, length, x, nimg=12):
    """sum_n exp(-alpha (x - n L)^2) and its x-derivative: the periodic
    Gaussian by direct image summation (Poisson dual of the theta series)."""
    n = np.arange(-nimg, nimg + 1)
    y = x[..., None] - n * np.asarray(length)[..., None]
    g = np.exp(-alpha * y * y)
    return g.sum(-1), (-2 * alpha * y * g).sum(-1)


def _sho_matrix(r1, r2, c, coshwt):
    # mat(i, j) = exp(-c ((|r2_i|^2 + |r1_j|^2) cosh - 2 r2_i . r1_j)), rows r2, columns r1
    a = (r2 * r2).sum(1)[:, None] + (r1 * r1).sum(1)[None, :]
    arg = -c * (a * coshwt - 2 * r2 @ r1.T)
    return np.exp(arg), arg


def _sho_distances(r1, r2, c, coshwt, mass, tau):
    mat, _ = _sho_matrix(r1, r2, c, coshwt)
    inv = np.linalg.inv(mat)
    n = len(r1)
    d1, d2 = np.zeros(n), np.zeros(n)
    for j in range(n):
        # grad_{r1_j} rho_ij = 2c (r2_i - cosh r1_j) rho_ij
        g = 2 * c * (r2 - coshwt * r1[j]) * mat[:, j][:, None]
        lg = (inv[j, :][:, None] * g).sum(0)
        d1[j] = math.sqrt(2 * mass / ((lg @ lg + 1e-15) * tau))
    for i in range(n):
        g = 2 * c * (r1 - coshwt * r2[i]) * mat[i, :][:, None]
        lg = (inv[:, i][:, None] * g).sum(0)
        d2[i] = math.sqrt(2 * mass / ((lg @ lg + 1e-15) * tau))
    return mat, d1, d2


def _free_distances(r1, r2, alpha, cell, mass, tau):
    n, D = r1.shape
    delta = r1[None, :, :] - r2[:, None, :]            # [i][j] = r1_j - r2_i
    val, grd = _image_sum_gaussian(alpha, np.asarray(cell[:D]), delta)
    mat = val.prod(-1)
    inv = np.linalg.inv(mat)
    # grad_{r1_j} rho_ij (d/dx of the periodic Gaussian in each dimension)
    g1 = np.stack([grd[..., k] * np.prod(np.delete(val, k, axis=-1), -1) for k in range(D)], -1)
    d1, d2 = np.zeros(n), np.zeros(n)
    for j in range(n):
        lg = (inv[j, :][:, None] * g1[:, j, :]).sum(0)
        d1[j] = math.sqrt(2 * mass / ((lg @ lg + 1e-15) * tau))
    for i in range(n):
        lg = (inv[:, i][:, None] * (-g1[i, :, :])).sum(0)
        d2[i] = math.sqrt(2 * mass / ((lg @ lg + 1e-15) * tau))
    return mat, d1, d2


@pytest.mark.parametrize("model", [capi.PQ_NODES_SHO, capi.PQ_NODES_FREE])
@pytest.mark.parametrize("ns", [3, 5, 8, 13, 20, 32])
def test_node_det_and_inverse_against_numpy(model, ns):
    """det = numpy.linalg.det of the closed-form matrix; the nodal distances
    d_j = sqrt(2m / ((|sum_i (rho^-1)_ji grad_j rho_ij|^2 + 1e-15) tau)) with
    numpy.linalg.inv.  Tolerances scale with the matrix condition number
    (partial-pivot LU is backward stable, so its forward error is eps * cond)."""
    D, M, tau, omega = 3, 64, 0.02, 1.0
    cell = 9.0
    d = capi.make_desc(D, M, 3, tau, [(ns, 1.0, -1.0), (2, 1.0, -1.0)], cell=cell,
                       sho=[(0, omega)], nodes=[(0, model, omega)], nsection=1)
    ch = ol.Chain(d)
    rng = np.random.default_rng(100 + ns + 7 * model)
    T = 1.0 / (tau * M)
    mass = 1.0
    checked = 0
    for trial in range(4):
        if model == capi.PQ_NODES_SHO:
            r1 = rng.uniform(-1.6, 1.6, size=(ns, D))
        else:
            r1 = rng.uniform(-0.5, 0.5, size=(ns, D)) * cell
        r2 = r1 + 0.25 * rng.standard_normal((ns, D))
        if model == capi.PQ_NODES_SHO:
            c = mass * 0.5 * omega / math.sinh(omega * 0.5 / T)
            mat, e1, e2 = _sho_distances(r1, r2, c, math.cosh(omega * 0.5 / T), mass, tau)
        else:
            mat, e1, e2 = _free_distances(r1, r2, mass * T, [cell] * 3, mass, tau)
        kappa = np.linalg.cond(mat)
        if kappa > 1e10:
            continue
        full1 = np.vstack([r1, np.zeros((2, D))])
        full2 = np.vstack([r2, np.zeros((2, D))])
        det, err = ch.node_evaluate(0, full1, full2, 0)
        assert err == 0
        want = np.linalg.det(mat)
        tol = max(1e-12, 64 * ns * np.finfo(float).eps * kappa)
        assert det == pytest.approx(want, rel=tol), (trial, det, want, kappa)
        d1, d2 = ch.node_distance(0, full1, full2, 0, ns)
        np.testing.assert_allclose(d1, e1, rtol=tol)
        np.testing.assert_allclose(d2, e2, rtol=tol)
        checked += 1
    assert checked >= 2


def test_node_det_sign_follows_permutation_parity():
    """Exchanging particles of one section permutes the matrix columns:
    det(P) = sign(P) det, for every permutation of 4 particles."""
    d = capi.make_desc(2, 64, 3, 0.02, [(4, 1.0, -1.0)], sho=[(0, 1.0)],
                       nodes=[(0, capi.PQ_NODES_SHO, 1.0)], nsection=1)
    ch = ol.Chain(d)
    rng = np.random.default_rng(41)
    r1 = rng.uniform(-1, 1, size=(4, 2))
    r2 = r1 + 0.2 * rng.standard_normal((4, 2))
    det0, _ = ch.node_evaluate(0, r1, r2, 0)
    for p in itertools.permutations(range(4)):
        inv = sum(1 for a in range(4) for b in range(a + 1, 4) if p[a] > p[b])
        det, _ = ch.node_evaluate(0, r1[list(p)], r2, 0)
        assert det == pytest.approx((-1) ** inv * det0, rel=1e-12)


# ------------------------------------------------------------------ g(r)

def _nearest_image_distance(a, b, cell, D):
    """|a - b| minimised over the images of b in a rectangular cell."""
    best = np.full(a.shape[:-1], np.inf)
    for shift in itertools.product((-1, 0, 1), repeat=D):
        v = a - b + np.array(shift) * np.asarray(cell[:D])
        best = np.minimum(best, np.sqrt((v * v).sum(-1)))
    return best


@pytest.mark.parametrize("D,cell", [(2, 3.0), (3, 2.5), (3, 1000.0)])
def test_pair_histogram_against_brute_force(D, cell):
    """Counts for species1 against each partner: ordered pairs i != j on the
    same slice, nearest-image distance, bins [min + k w, min + (k+1) w),
    r < min and r >= max dropped; a same-species histogram counts each pair
    twice."""
    M, counts_ = 24, (5, 3, 4)
    d = capi.make_desc(D, M, 2, 0.05, [(n, 1.0, -1.0) for n in counts_], cell=cell, nsection=1)
    ch = ol.Chain(d)
    rng = np.random.default_rng(7 * D)
    span = min(cell, 4.0)
    beads = rng.uniform(-0.5, 0.5, size=(M, sum(counts_), D)) * span
    ch.set_beads(np.ascontiguousarray(beads))
    first = np.cumsum((0,) + counts_)
    rmin, rmax, nbin = 0.3, 1.1 * span / 2, 37
    edges = np.linspace(rmin, rmax, nbin + 1)
    for s1, s2 in ((0, 1), (0, 0), (2, 1), (1, 2)):
        got = ch.pair_histogram(s1, s2, rmin, rmax, nbin)
        ii = range(first[s1], first[s1 + 1])
        jj = range(first[s2], first[s2 + 1])
        dist = np.array([_nearest_image_distance(beads[:, i], beads[:, j], [cell] * 3, D)
                         for i in ii for j in jj if i != j]).ravel()
        # no distance may sit within rounding of a bin edge, or the two binnings could differ
        frac = (dist - rmin) / (rmax - rmin) * nbin
        assert np.min(np.abs(frac - np.rint(frac))) > 1e-9
        want, _ = np.histogram(dist[(dist >= rmin) & (dist < rmax)], bins=edges)
        assert np.array_equal(got, want.astype(np.uint64)), (s1, s2)
        assert int(got.sum()) == int(((dist >= rmin) & (dist < rmax)).sum())
        if s1 == s2:
            assert all(int(v) % 2 == 0 for v in got)


def test_pair_cf_partner_list_sums_partners():
    """PairCFEstimator's species2..n partner list: the counts are the sum of
    the single-partner histograms, a partner listed twice counting twice."""
    d = capi.make_desc(2, 16, 2, 0.05, [(3, 1.0, -1.0), (2, 1.0, -1.0), (4, 1.0, 1.0)], cell=5.0,
                       nsection=1)
    ch = ol.Chain(d)
    rng = np.random.default_rng(3)
    ch.set_beads(np.ascontiguousarray(rng.uniform(-2, 2, size=(16, 9, 2))))
    args = (0.0, 2.5, 20)
    h01, h00, h02 = (ch.pair_histogram(0, s, *args) for s in (1, 0, 2))
    assert np.array_equal(ch.pair_cf_histogram([0, 1, 2], *args), h01 + h02)
    assert np.array_equal(ch.pair_cf_histogram([0, 0, 1], *args), h00 + h01)
    assert np.array_equal(ch.pair_cf_histogram([0, 2, 2], *args), 2 * h02)
    assert not ch.pair_cf_histogram([0], *args).any()


# --------------------------------------------------- FreeMoverPBC on a torus

def _torus_desc(nlevel, nsection):
    # one free particle (m = 1, lambda = 1/2) on a 2D torus of side 1 at beta = 0.2,
    # M = 128: L^2 / (4 lambda beta) = 2.5, so windings +-1 carry 14% of the weight
    return capi.make_desc(2, 128, nlevel, 0.2 / 128, [(1, 1.0, 0.0)], cell=1.0,
                          mover=capi.PQ_MOVER_FREE_PBC, nsection=nsection)


def torus_exact(L=1.0, lam=0.5, beta=0.2, D=2):
    """<w^2> per dimension and the energy D(1/(2 beta) - <w^2> L^2/(4 lambda beta^2))
    of a free particle on a torus: Z_d = sum_w exp(-w^2 L^2/(4 lambda beta))
    (4 pi lambda beta)^(-1/2) L, E = -d ln Z/d beta."""
    w = np.arange(-8, 9)
    p = np.exp(-w * w * L * L / (4 * lam * beta))
    w2 = float((w * w * p).sum() / p.sum())
    return w2, D * (1 / (2 * beta) - w2 * L * L / (4 * lam * beta * beta))


def windings(beads, L=1.0):
    """Winding number per dimension of a closed single-particle path [M][1][D]."""
    b = beads[:, 0, :]
    dl = np.diff(np.concatenate([b, b[:1]]), axis=0)
    dl -= L * np.round(dl / L)
    return np.rint(dl.sum(0) / L)


def test_free_mover_pbc_torus_windings_and_energy():
    """Sections of the whole path (2^L = M) whose top levels sample through
    periodic Gaussians (fewer than 16 terms) and the image choice of
    FreeMoverPBC: the winding-number law and the thermal energy of the torus.
    The sampled energy at these sizes: 1.44 +- 0.07 over 5e5 measurements
    against 1.467; the classical value without the torus is 10."""
    w2_exact, e_exact = torus_exact()
    ch = ol.Chain(_torus_desc(7, 1))
    ch.set_beads(np.zeros((128, 1, 2)))
    ch.rng_philox(11, 0)
    ch.run_sweeps(1000)
    ws = []
    for _ in range(12000):
        ch.run_sweeps(1)
        ws.append(windings(ch.get_beads()))
    ws = np.array(ws)
    w2 = float((ws * ws).mean())
    # block error of <w^2> over 24 blocks
    blocks = (ws * ws).reshape(24, -1).mean(1)
    err = blocks.std(ddof=1) / math.sqrt(len(blocks))
    assert abs(w2 - w2_exact) < max(5 * err, 0.01), (w2, w2_exact, err)
    frac1 = float(np.mean(np.abs(ws) == 1))
    assert abs(frac1 - 2 * math.exp(-2.5) / (1 + 2 * math.exp(-2.5))) < 0.02
    e = ch.energy_blocks(20, 4000)
    eerr = e.std(ddof=1) / math.sqrt(len(e))
    assert abs(e.mean() - e_exact) < max(4 * eerr, 0.35), (e.mean(), eerr, e_exact)


def test_sho_node_matrix_rank_one_at_low_temperature():
    """Why the device parity tests of the 3D and 2D batch chains set a node
    temperature: at the simulation temperature of a chain with M = 2048, tau
    = 0.05 (beta/2 = 51.2), 2c r.r' ~ 1e-22 is below FP64 resolution next to
    c cosh ~ 0.5, so the closed-form SHO matrix is rank one to working
    precision and its determinant is rounding noise -- in numpy as in the
    oracle (and the reference).  At temperature 0.5 the same positions give a
    resolved determinant that numpy and the oracle agree on."""
    d = capi.batch3d_descs(4096)[114]                 # N_s 20 + 20, M 2048
    assert d.nslice == 2048 and d.species[0].count == 20
    ns = 20
    rng = np.random.default_rng(1)
    r1 = rng.standard_normal((d.npart, 3))
    r2 = r1 + 0.3 * rng.standard_normal((d.npart, 3))
    for T, resolved in ((1.0 / (d.tau * d.nslice), False), (0.5, True)):
        c = 0.5 / math.sinh(0.5 / T)
        mat, _ = _sho_matrix(r1[:ns], r2[:ns], c, math.cosh(0.5 / T))
        sv = np.linalg.svd(mat, compute_uv=False)
        e = capi.ChainDesc()
        capi.C.pointer(e)[0] = d
        e.nodes[0].temperature = T
        det, err = ol.Chain(e).node_evaluate(0, r1, r2, 0)
        assert err == 0
        if resolved:
            assert sv[-1] / sv[0] > 1e-8
            assert det == pytest.approx(np.linalg.det(mat), rel=1e-9)
        else:
            assert sv[1] / sv[0] < 1e-14
            # |det| below eps^(N_s - 1) of the diagonal product: no digit is significant
            assert abs(det) < np.prod(np.abs(np.diag(mat))) * 1e-15 ** (ns - 1) * 1e30
