"""The oracle's FreeParticleNodes::MatrixUpdate (useUpdates, FreeParticleNodes.cc:
894-1058, FixedNodeAction.cc:61-92 and 366) pinned on the CPU against numpy:

* evaluateChange: ratio * dmValue equals numpy's determinant of the moved
  slice's new matrix, for one and for two moving particles of a species, and
  after accepted moves (the updated inverses stay right);
* evaluateNewInverse + evaluateNewDistance: the level-0 nodal action equals a
  numpy restatement with numpy inverses of the new matrices, with the
  update path's own distance formula (no scale factor; in the first loop each
  gradient component times its own dimension's value only), which differs
  from evaluateDistance's, and that difference is checked to be there;
* without the nodal action, acceptLastMove's in-place row updates.
The periodic Gaussians are image sums here, the oracle's the theta series."""
import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi

L_BOX = 4.0


def _desc(updates, nodal=True, counts=(4, 3), nslice=64, nlevel=3):
    d = capi.make_desc(2, nslice, nlevel, 0.05, [(counts[0], 1.0, -1.0), (counts[1], 1.5, -1.0)],
                       cell=L_BOX,
                       nodes=[(0, capi.PQ_NODES_FREE, 0.0, updates), (1, capi.PQ_NODES_FREE, 0.0, updates)],
                       mover=capi.PQ_MOVER_FREE_PBC)
    for a in range(d.nnodes):
        d.nodes[a].nodal_action = 1 if nodal else 0
    return d


def _pg(alpha, x):
    """sum over images of exp(-alpha (x - m L)^2) and its derivative"""
    m = np.arange(-8, 9)
    y = x[..., None] - m * L_BOX
    e = np.exp(-alpha * y * y)
    return e.sum(-1), (-2 * alpha * y * e).sum(-1)


def _min_image(d):
    return d - L_BOX * np.round(d / L_BOX)


def _matrix(desc, sp, r1, r2):
    """A[i][j] = prod_d PG(r1_j - r2_i) of species sp (FreeParticleNodes.cc:79-90)"""
    f = desc.ifirst(sp) if hasattr(desc, "ifirst") else sum(desc.species[k].count for k in range(sp))
    n = desc.species[sp].count
    alpha = desc.species[sp].mass / (desc.tau * desc.nslice)
    delta = _min_image(r1[None, f:f + n, :] - r2[f:f + n, None, :])
    v, _ = _pg(alpha, delta)
    return np.prod(v, axis=-1)


def _distances(desc, sp, r1, r2, binv, scale, own_dim_first):
    """evaluateDistance (own_dim_first False, with scale) or the update path's
    evaluateNewDistance (own_dim_first True, no scale), both [2][n]"""
    f = sum(desc.species[k].count for k in range(sp))
    n = desc.species[sp].count
    mass = desc.species[sp].mass
    alpha = mass / (desc.tau * desc.nslice)
    out = np.zeros((2, n))
    for p in range(2):
        for a in range(n):
            lg = np.zeros(2)
            for b in range(n):
                jp, ip = (a, b) if p == 0 else (b, a)
                delta = r1[f + jp] - r2[f + ip] if p == 0 else r2[f + ip] - r1[f + jp]
                delta = _min_image(delta)
                v, g = _pg(alpha, delta)
                grad = g / (v + 1e-300)
                if own_dim_first and p == 0:
                    grad = grad * v
                else:
                    grad = grad * np.prod(v)
                lg += binv[jp, ip] * grad * (1.0 if own_dim_first else scale)
            out[p, a] = np.sqrt(2 * mass / ((lg @ lg + 1e-15) * desc.tau))
    return out


def _sections(desc, seed, gap=None, noise=0.12):
    """both sections near one lattice configuration per species (spacing 2 in
    the box of 4), so every slice's determinant has slice 0's sign; with gap,
    particle 1 sits that far from particle 0, near their node"""
    rng = np.random.default_rng(seed)
    S = (1 << desc.nlevel) + 1
    base = np.zeros((1, desc.npart, 2))
    i = 0
    for sp in range(desc.nspecies):
        for k in range(desc.species[sp].count):
            base[0, i] = [2.0 * (k % 2) - 1.0 + 0.5 * sp, 2.0 * ((k // 2) % 2) - 1.0 + 0.5 * sp]
            base[0, i] += 0.3 * (k // 4)
            i += 1
    if gap is not None:
        base[0, 1] = base[0, 0] + [gap, 0.0]
    sec1 = base + noise * rng.standard_normal((S, desc.npart, 2))
    sec2 = base + noise * rng.standard_normal((S, desc.npart, 2))
    return np.ascontiguousarray(sec1), np.ascontiguousarray(sec2)


def _move(sec1, idx, rng, step=0.08):
    mov = np.ascontiguousarray(sec1[:, idx, :].copy())
    mov[1:-1] += step * rng.standard_normal(mov[1:-1].shape)
    return mov


def _species_of(desc, i):
    return 0 if i < desc.species[0].count else 1


@pytest.mark.parametrize("idx", [[2], [0, 3], [1, 5]])
def test_evaluate_change_equals_the_new_determinant(idx):
    """newDMValue = evaluateChange * dmValue against numpy's det of the new
    matrix on every slice of every level, over three accepted moves."""
    d = _desc(True, nodal=False)
    oc = ol.Chain(d)
    sec1, sec2 = _sections(d, 5)
    oc.fn_initialize(sec1, sec2)
    rng = np.random.default_rng(11)
    S = (1 << d.nlevel) + 1
    for move in range(3):
        mov = _move(sec1, idx, rng)
        for level in range(d.nlevel, -1, -1):
            dA = oc.fn_action_difference(mov, idx, level)
            assert dA == 0.0, (move, level, dA)
        new1 = sec1.copy()
        new1[:, idx, :] = mov
        for sp in sorted({_species_of(d, i) for i in idx}):
            ndm, _ = oc.fn_state(sp, new=True)
            for t in range(1, S):
                want = np.linalg.det(_matrix(d, sp, new1[t], sec2[t]))
                assert abs(ndm[t] - want) <= 1e-10 * abs(want), (move, sp, t, ndm[t], want)
        oc.fn_accept()
        sec1 = new1


def test_level_zero_nodal_action_of_the_update_path():
    """the level-0 nodal action with useUpdates against numpy inverses of the
    new matrices and evaluateNewDistance's formula; the same move without
    updates (evaluateDistance) gives a different action in 2D."""
    d = _desc(True)
    oc = ol.Chain(d)
    sec1, sec2 = _sections(d, 9, gap=0.35, noise=0.04)
    oc.fn_initialize(sec1, sec2)
    assert oc.node_scale(0) == 1.0 and oc.node_scale(1) == 1.0
    rng = np.random.default_rng(4)
    S = (1 << d.nlevel) + 1
    idx = [1]
    mov = _move(sec1, idx, rng, step=0.03)
    for level in range(d.nlevel, 0, -1):
        assert oc.fn_action_difference(mov, idx, level) == 0.0
    got = oc.fn_action_difference(mov, idx, 0)
    new1 = sec1.copy()
    new1[:, idx, :] = mov
    sp = 0
    od = np.zeros((S, 2, d.species[sp].count))
    nd = np.zeros_like(od)
    for t in range(S):
        od[t] = _distances(d, sp, sec1[t], sec2[t], np.linalg.inv(_matrix(d, sp, sec1[t], sec2[t])), 1.0, False)
        if t:
            nd[t] = _distances(d, sp, new1[t], sec2[t], np.linalg.inv(_matrix(d, sp, new1[t], sec2[t])), 1.0, True)
    nd[0] = od[0]
    want = 0.0
    for t in range(1, S):
        want += np.sum(np.log((1 - np.exp(-od[t] * od[t - 1])) / (1 - np.exp(-nd[t] * nd[t - 1]))))
    assert abs(want) > 1e-4, want   # the particles are near their node
    assert abs(got - want) <= 1e-9 * max(1.0, abs(want)), (got, want)
    _, o_dist = oc.fn_state(sp)
    assert np.allclose(o_dist, od, rtol=1e-10, atol=0)
    # evaluateDistance on the same move: another action
    plain = ol.Chain(_desc(False))
    plain.fn_initialize(sec1, sec2)
    for level in range(d.nlevel, 0, -1):
        plain.fn_action_difference(mov, idx, level)
    other = plain.fn_action_difference(mov, idx, 0)
    assert abs(other - want) > 1e-6 * max(1.0, abs(want)), (other, want)


def test_accept_updates_the_inverses_in_place_without_nodal_action():
    """acceptLastMove's in-place branch (no level-0 evaluateNewInverse): the
    determinants of a later move still equal numpy's, the moved particle
    moving again."""
    d = _desc(True, nodal=False, counts=(5, 2))
    oc = ol.Chain(d)
    sec1, sec2 = _sections(d, 21)
    oc.fn_initialize(sec1, sec2)
    rng = np.random.default_rng(8)
    S = (1 << d.nlevel) + 1
    for move, idx in enumerate(([3], [3], [0, 3], [4])):
        mov = _move(sec1, idx, rng, step=0.1)
        for level in range(d.nlevel, -1, -1):
            assert oc.fn_action_difference(mov, idx, level) == 0.0
        new1 = sec1.copy()
        new1[:, idx, :] = mov
        ndm, _ = oc.fn_state(0, new=True)
        for t in range(1, S - 1):
            want = np.linalg.det(_matrix(d, 0, new1[t], sec2[t]))
            assert abs(ndm[t] - want) <= 1e-10 * abs(want), (move, t)
        oc.fn_accept()
        sec1 = new1


def test_use_updates_is_refused_for_sho_nodes():
    d = capi.make_desc(2, 32, 2, 0.05, [(3, 1.0, -1.0)], sho=[(0, 1.0)],
                       nodes=[(0, capi.PQ_NODES_SHO, 1.0, True)])
    lib = capi.load_library()
    arr = capi.descs_array([d])
    h = capi.C.c_void_p()
    assert lib.pq_batch_create(arr, 1, capi.PQ_PREC_FP64, 0, capi.C.byref(h)) == capi.PQ_ERR_ARG
