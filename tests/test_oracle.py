"""The oracle pinned before it is trusted: reference known-answer tests, the
reference's own Coulomb/PeriodicGaussian code (oracle/_ref, when built), libc
rand(), Random123's Philox vectors, closed forms and invariants chosen so that
a dropped term, a wrong sign or a wrong index fails one of them."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as ol
import piqmc_capi as capi


@pytest.fixture(scope="module")
def L():
    return ol.load()


# ------------------------------------------------ Coulomb link action KATs

@pytest.mark.parametrize("stau,expect", [(0.0, 0.0), (0.01, 0.017717130566979017),
                                          (0.1, 0.17650980432591801), (-0.01, -0.017731958122632733),
                                          (-0.1, -0.17799263565884202)])
def test_coulomb1d_origin(L, stau, expect):
    # unit-tests/action/coulomb/Coulomb1DLinkActionTest.cpp:20-36
    assert L.oq_coulomb1d(stau, -1, 0.0) == pytest.approx(expect, rel=1e-15, abs=1e-300)


@pytest.mark.parametrize("which,expect", [(0, 0.010025994988629595), (1, 0.00085593234703468702),
                                          (2, 0.00012817165840892054), (3, 2.3263720535382824e-5),
                                          (4, 4.3433495124893814e-6)])
def test_coulomb1d_orders(L, which, expect):
    # Coulomb1DLinkActionTest.cpp:38-77 (stau 0.1, reff 1)
    assert L.oq_coulomb1d(0.1, which, 1.0) == pytest.approx(expect, rel=4e-16)


@pytest.mark.parametrize("which,reff,expect", [(0, 0.0, 0.017717130566979017),
                                               (0, 1.0, 9.9843305484870431e-5),
                                               (1, 1.0, 8.5031825295702207e-6),
                                               (2, 1.0, 1.3429133944492796e-6),
                                               (3, 1.0, 1.637776980179888e-7),
                                               (4, 1.0, 4.0165622775654946e-8)])
def test_coulomb3d_orders(L, which, reff, expect):
    # unit-tests/action/coulomb/Coulomb3DLinkActionTest.cpp:32-67 (stau 0.01)
    assert L.oq_coulomb3d(0.01, which, reff) == pytest.approx(expect, rel=4e-16)


def _link_action(L, d1, d2, q1q2=-1.0, eps=12.0, mu=1.0, dtau=0.1):
    """CoulombLinkAction::getValue (src/action/coulomb/CoulombLinkAction.cpp:16-44), norder 3."""
    r = 0.5 * (np.linalg.norm(d1) + np.linalg.norm(d2))
    s2 = np.dot(d1 - d2, d1 - d2) / (r * r + 1e-200)
    stau = q1q2 / eps * math.sqrt(2 * mu * dtau)
    reff = 2.0 * mu * q1q2 * r / eps
    u = L.oq_coulomb3d(stau, 0, reff)
    for k in range(1, 4):
        u += L.oq_coulomb3d(stau, k, reff) * s2 ** k
    return u


def test_coulomb_link_action_kats(L):
    # unit-tests/action/coulomb/CoulombLinkActionTest.cpp:105-128
    z = np.zeros(3)
    assert _link_action(L, z, z) == pytest.approx(-0.066158726800149281, rel=1e-14)
    d = np.array([5.0, 0, 0])
    assert _link_action(L, d, d) == pytest.approx(-0.0016612138747140238, rel=1e-14)


def test_coulomb_bloch_equation(L):
    """The tabulated link action solves the Bloch equation (CoulombLinkActionTest.cpp:142-153):
    -lap(rho) + V rho = -d rho / d tau, with rho = rho_free * exp(-u)."""
    PI = 3.141592653389793
    q1q2, eps, mu, dtau = -1.0, 12.0, 1.0, 0.1

    def prop(d1, d2, t):
        r2 = np.dot(d1 - d2, d1 - d2)
        free = math.exp(-0.25 * r2 / t) * (4 * PI * t) ** (-1.5)
        return free * math.exp(-_link_action(L, d1, d2, q1q2, eps, mu, t))

    d1 = np.array([5.0, 0, 0])
    d2 = d1.copy()
    h = 1e-5
    lap = -6 * prop(d1, d2, dtau)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        lap += prop(d1 + e, d2, dtau) + prop(d1 - e, d2, dtau)
    lap /= h * h
    dt = 1e-10
    udot = (prop(d1, d2, dtau + dt) - prop(d1, d2, dtau - dt)) / (2 * dt)
    lhs = -lap + q1q2 / (5.0 * eps) * prop(d1, d2, dtau)
    assert abs(lhs + udot) < 5e-5


def test_coulomb_against_reference_build(L):
    R = ol.load_ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(0)
    for _ in range(200):
        stau = float(rng.uniform(-0.5, 0.5)) or 0.1
        reff = float(rng.uniform(-20, 20))
        for w in range(-1, 5):
            assert L.oq_coulomb1d(stau, w, reff) == R.ref_coulomb1d(stau, w, reff)
        for w in range(5):
            assert L.oq_coulomb3d(stau, w, reff) == R.ref_coulomb3d(stau, w, reff)


# ------------------------------------------------------ PeriodicGaussian

def test_periodic_gaussian_kats(L):
    # unit-tests/util/PeriodicGaussianTest.cc:26-91 (alpha 2.5, length 1.5)
    assert L.oq_pg_nterms(2.5, 1.5) == 5
    assert L.oq_pg_nterms(2.5, 20.0) == 64
    out = np.zeros(3)
    cases = [(0.0, 1.0072131266104112, 0.0, -4.6303272041148782),
             (0.4999 * 1.5, 0.49012750361989998, -0.00066637454689608252, 4.442496431737875),
             (-0.3 * 1.5, 0.66635569537953232, 1.0233852136355668, 1.4777706970511808)]
    for x, v, g, d2 in cases:
        L.oq_pg_eval(2.5, 1.5, x, ol.dptr(out))
        assert out[0] == pytest.approx(v, rel=4e-16)
        assert out[1] == pytest.approx(g, rel=4e-15, abs=1e-300)
        assert out[2] == pytest.approx(d2, rel=4e-15)


def test_periodic_gaussian_is_image_sum(L):
    """Poisson summation: the theta series equals sum_m exp(-alpha (x - m L)^2)."""
    out = np.zeros(3)
    for alpha, length in ((2.5, 1.5), (0.05, 10.0), (0.3, 7.0)):
        for x in np.linspace(-0.5 * length, 0.5 * length, 7):
            L.oq_pg_eval(alpha, length, x, ol.dptr(out))
            m = np.arange(-60, 61)
            brute = np.exp(-alpha * (x - m * length) ** 2).sum()
            assert out[0] == pytest.approx(brute, rel=1e-13)


def test_periodic_gaussian_against_reference_build(L):
    R = ol.load_ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    a, b = np.zeros(3), np.zeros(3)
    rng = np.random.default_rng(1)
    for _ in range(300):
        alpha, length = float(rng.uniform(0.01, 5)), float(rng.uniform(0.5, 20))
        x = float(rng.uniform(-length, length))
        L.oq_pg_eval(alpha, length, x, ol.dptr(a))
        R.ref_pg_eval(alpha, length, x, ol.dptr(b))
        assert L.oq_pg_nterms(alpha, length) == R.ref_pg_nterms(alpha, length)
        assert np.array_equal(a, b)


# ------------------------------------------------------------ SuperCell

def test_supercell_pbc_kat(L):
    # unit-tests/util/SuperCellTest.cc:20-32
    v = np.array([0.1, 2.0, 2.5])
    L.oq_pbc(ol.dptr(np.array([1.0, 2.0, 4.0])), 3, ol.dptr(v))
    assert list(v) == [0.1, 0.0, -1.5]


def test_pbc_is_minimum_image(L):
    rng = np.random.default_rng(2)
    a = np.array([3.0, 4.0, 5.0])
    for _ in range(500):
        v = rng.uniform(-12, 12, size=3)
        w = v.copy()
        L.oq_pbc(ol.dptr(a), 3, ol.dptr(w))
        imgs = [v - a * np.array(k) for k in np.ndindex(9, 9, 9)]
        imgs = [v - a * (np.array(k) - 4) for k in np.ndindex(9, 9, 9)]
        best = min(imgs, key=lambda z: np.dot(z, z))
        if np.dot(v, v) > 0.25 * 9:   # pbc only applies beyond half the shortest side
            assert np.allclose(w, best, atol=1e-12)


# ------------------------------------------------------------- uniforms

def test_glibc_rand_clone_matches_libc(L):
    libc = C.CDLL("libc.so.6")
    for seed in (0, 1, 4, 149, 12345):
        libc.srand(seed)
        want = [libc.rand() for _ in range(1000)]
        got = (C.c_int32 * 1000)()
        L.oq_glibc_sequence(seed, 1000, got)
        assert list(got) == want


def test_philox_known_answer(L):
    # Random123 kat_vectors: philox4x32 10, counter 0, key 0
    out = (C.c_uint32 * 4)()
    L.oq_philox_raw((C.c_uint32 * 4)(0, 0, 0, 0), (C.c_uint32 * 2)(0, 0), out)
    assert list(out) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert [L.oq_philox_u32(0, 0, k) for k in range(4)] == list(out)


# --------------------------------------------------------------- actions

def test_sho_action_kat():
    """unit-tests/action/SHOActionTest.cc:58-69: bead at slice 32 of 64 moved to
    (1,2,3), tau 0.1, omega = m = 1, level 0."""
    d = capi.make_desc(3, 128, 6, 0.1, [(1, 1.0, 0.0)], sho=[(0, 1.0)], spring=False)
    ch = ol.Chain(d)
    sec = np.zeros((65, 1, 3))
    mov = np.zeros((65, 1, 3))
    assert ch.action_difference(sec, mov, [0], 0, mask=2) == 0.0
    mov[32, 0] = [1, 2, 3]
    assert ch.action_difference(sec, mov, [0], 0, mask=2) == pytest.approx(0.46635586, rel=2.4e-7)


def test_pair_action_on_grid_points():
    """Particle 1 sits at the origin; particle 0 moves on a circle whose radius
    is a grid knot r_a, so q = r_a and the lookup returns the knot values (the
    interpolation weight is 0 or 1).  The old path keeps angle 0 (s2 = 0), the
    new one turns by theta per slice (s2 = 2(1-cos(theta s))), so
    dS = sum_links s [sum_k u_k(r_a) s2^k - u_0(r_a)] below level 3.  From level 3
    on only u_0(q) enters, so the turning path costs nothing there.  This pins the
    grid index, the Horner series in s2, the stride factor and the level cut."""
    d = capi.make_desc(3, 64, 3, 0.05, [(1, 1.0, -1.0), (1, 1.0, -1.0)], spring=False,
                       coulomb=dict(epsilon=1.0, norder=3, rmin=0.01, rmax=20.0, ngpts=300))
    ch = ol.Chain(d)
    meta, grid, u, ut = ch.pair_table(0)
    rgrid = (1.0 / grid[0]) * np.exp(np.arange(meta[3]) / grid[1])
    S = 9

    def circle(r, theta):
        return np.array([[r * math.cos(theta * t), r * math.sin(theta * t), 0.0] for t in range(S)])

    for level, ia, ib, theta in [(0, 40, 40, 0.3), (1, 200, 200, 0.05), (2, 77, 77, 0.2),
                                 (3, 150, 150, 0.2)]:
        s = 1 << level
        sec = np.zeros((S, 2, 3))
        sec[:, 0] = circle(rgrid[ib], 0.0)
        mov = circle(rgrid[ia], theta)[:, None, :].copy()
        got = ch.action_difference(sec, mov, [0], level, mask=4)
        nl = (S - 1) // s
        if level < 3:
            s2 = 2 * (1 - math.cos(theta * s))
            want = nl * s * (sum(u[ia][k] * s2 ** k for k in range(4)) - u[ia][0])
        else:
            want = 0.0
        assert got == pytest.approx(want, rel=1e-9, abs=1e-13), (level, got, want)


def test_action_difference_antisymmetric():
    """Swapping the old and new paths of the moving particle negates dS."""
    d = capi.dot3d_2e_desc()
    ch = ol.Chain(d)
    rng = np.random.default_rng(4)
    S = 9
    for level in range(4):
        sec = rng.standard_normal((S, 2, 3))
        mov = sec[:, 0:1, :].copy()
        mov[1:-1] += 0.4 * rng.standard_normal((S - 2, 1, 3))
        a = ch.action_difference(sec, mov, [0], level)
        sec2 = sec.copy()
        sec2[:, 0] = mov[:, 0]
        b = ch.action_difference(sec2, sec[:, 0:1, :].copy(), [0], level)
        assert a == pytest.approx(-b, rel=1e-12, abs=1e-13)


# ------------------------------------------------------------ node models

@pytest.mark.parametrize("model", [capi.PQ_NODES_SHO, capi.PQ_NODES_FREE])
def test_node_det_and_distance_pinned(model):
    """Row-swap antisymmetry of det, and the nodal distance against a
    finite-difference gradient of log|det|: d_j = sqrt(2m/((|grad_j
    log|det||^2 + 1e-15) tau)).  The determinant values themselves are checked
    against numpy.linalg.det of the closed-form matrix in test_oracle_pins.py."""
    d = capi.make_desc(2, 64, 4, 0.02, [(3, 1.0, -1.0)], cell=6.0, sho=[(0, 1.0)],
                       nodes=[(0, model, 1.0)])
    ch = ol.Chain(d)
    rng = np.random.default_rng(8)
    r1 = rng.uniform(-1.5, 1.5, size=(3, 2))
    r2 = r1 + 0.3 * rng.standard_normal((3, 2))
    det, err = ch.node_evaluate(0, r1, r2, 0)
    assert err == 0
    sw = r1.copy()
    sw[[0, 1]] = sw[[1, 0]]
    det_sw, _ = ch.node_evaluate(0, sw, r2, 0)
    assert det_sw == pytest.approx(-det, rel=1e-12)
    ch.node_evaluate(0, r1, r2, 0)
    d1, d2 = ch.node_distance(0, r1, r2, 0, 3)
    tau = d.tau
    h = 1e-6
    for which, rr, dd in ((0, r1, d1), (1, r2, d2)):
        for j in range(3):
            g = np.zeros(2)
            for k in range(2):
                p = rr.copy(); p[j, k] += h
                m = rr.copy(); m[j, k] -= h
                a1, a2 = (p, r2) if which == 0 else (r1, p)
                b1, b2 = (m, r2) if which == 0 else (r1, m)
                dp, _ = ch.node_evaluate(0, a1, a2, 0)
                dm, _ = ch.node_evaluate(0, b1, b2, 0)
                g[k] = (math.log(abs(dp)) - math.log(abs(dm))) / (2 * h)
            want = math.sqrt(2 * 1.0 / ((np.dot(g, g) + 1e-15) * tau))
            assert dd[j] == pytest.approx(want, rel=1e-6)


def test_sho_node_matrix_single_particle():
    """N_s = 1: the determinant is the SHO density matrix at beta/2 itself."""
    d = capi.make_desc(3, 100, 3, 0.04, [(1, 1.0, -1.0)], sho=[(0, 0.7)],
                       nodes=[(0, capi.PQ_NODES_SHO, 0.7)])
    ch = ol.Chain(d)
    r1 = np.array([[0.3, -0.2, 0.5]])
    r2 = np.array([[-0.1, 0.4, 0.2]])
    det, _ = ch.node_evaluate(0, r1, r2, 0)
    beta_half = 0.5 * d.tau * d.nslice
    w = 0.7
    expect = math.exp(-w / (2 * math.sinh(w * beta_half))
                      * ((r1 @ r1.T + r2 @ r2.T)[0, 0] * math.cosh(w * beta_half) - 2 * (r1 @ r2.T)[0, 0]))
    assert det == pytest.approx(expect, rel=1e-13)


def test_fixed_node_reverse_move_negates():
    """dS(A->B) = -dS(B->A) for the nodal action at level 0, and a crossing gives 2e100."""
    d = capi.make_desc(2, 64, 2, 0.02, [(3, 1.0, -1.0)], sho=[(0, 1.0)],
                       nodes=[(0, capi.PQ_NODES_SHO, 1.0)])
    ch = ol.Chain(d)
    rng = np.random.default_rng(12)
    S = 5
    base = rng.uniform(-1, 1, size=(1, 3, 2))
    secA = base + 0.05 * rng.standard_normal((S, 3, 2))
    sec2 = base + 0.05 * rng.standard_normal((S, 3, 2))
    movB = secA[:, 1:2, :].copy()
    movB[1:-1] += 0.05 * rng.standard_normal((S - 2, 1, 2))
    ch.fn_initialize(secA, sec2)
    for lv in (2, 1):
        ch.fn_action_difference(movB, [1], lv)
    ab = ch.fn_action_difference(movB, [1], 0)
    secB = secA.copy()
    secB[:, 1] = movB[:, 0]
    ch.fn_initialize(secB, sec2)
    for lv in (2, 1):
        ch.fn_action_difference(secA[:, 1:2, :].copy(), [1], lv)
    ba = ch.fn_action_difference(secA[:, 1:2, :].copy(), [1], 0)
    assert ab < 1e99 and ab == pytest.approx(-ba, rel=1e-10, abs=1e-12)
    # swap particle 1 onto particle 0's path at slice 2: the determinant changes sign
    movX = secA[:, 1:2, :].copy()
    movX[2, 0] = secA[2, 0] + 1e-3
    movX[1, 0] = 0.5 * (secA[0, 1] + movX[2, 0])
    movX[3, 0] = 0.5 * (secA[4, 1] + movX[2, 0])
    ch.fn_initialize(secA, sec2)
    r = ch.fn_action_difference(movX, [1], 1)
    assert r == 2e100


# ---------------------------------------------------------------- sampler

def test_sho_thermal_energy():
    """test/system/sho: 1 particle, 3D, omega 0.5, T 1, 32 slices, nlevel 5:
    E = 0.75 coth(0.25) within 0.08 (test/system/sho/test_sho.py:27-32)."""
    d = capi.make_desc(3, 32, 5, 1 / 32, [(1, 1.0, -1.0)], cell=20.0, sho=[(0, 0.5)], nsection=1)
    ch = ol.Chain(d)
    ch.set_beads(np.zeros((32, 1, 3)))
    ch.rng_glibc(0)
    ch.run_sweeps(100)
    blocks = ch.energy_blocks(500, 200)
    assert blocks.mean() == pytest.approx(0.75 / math.tanh(0.25), abs=0.08)


def test_two_fermion_sho_fixed_node_energy():
    """Two spin-polarised non-interacting fermions in a 2D dot with exact SHO
    nodes: E = -d ln Z_F / d beta, Z_F = (Z1(b)^2 - Z1(2b))/2, Z1 = (2 sinh(b/2))^-2."""
    beta, M = 2.0, 64
    d = capi.make_desc(2, M, 3, beta / M, [(2, 1.0, -1.0)], sho=[(0, 1.0)],
                       nodes=[(0, capi.PQ_NODES_SHO, 1.0)], nsection=M // 8)
    ch = ol.Chain(d)
    ch.set_beads(np.broadcast_to(np.array([[0.7, 0.0], [-0.7, 0.1]]), (M, 2, 2)).copy())
    ch.rng_philox(5, 0)
    ch.run_sweeps(300)
    blocks = ch.energy_blocks(40, 100)

    def lnz(b):
        z1 = lambda x: (2 * math.sinh(x / 2)) ** -2
        return math.log(0.5 * (z1(b) ** 2 - z1(2 * b)))
    h = 1e-5
    exact = -(lnz(beta + h) - lnz(beta - h)) / (2 * h)
    err = blocks.std() / math.sqrt(len(blocks))
    assert abs(blocks.mean() - exact) < max(5 * err, 0.15), (blocks.mean(), exact, err)


def test_decisions_repeatable_and_counted():
    d = capi.dot3d_2e_desc()
    beads = capi.initial_beads(d, 1)
    a, b = ol.Chain(d), ol.Chain(d)
    for ch in (a, b):
        ch.set_beads(beads)
        ch.rng_glibc(1)
        ch.run_sweeps(3)
    sa, sb = a.stats(), b.stats()
    assert list(sa.accepts) == list(sb.accepts) and sa.draws == sb.draws
    # every section draws once, every sample once more, every tried level once
    tries = sum(sa.tries)
    assert sa.tries[3] == 3 * d.nsection * 2
    assert sa.bead_updates == sum(sa.tries[l] << (2 - l) for l in range(3))
    assert sa.draws >= 3 * d.nsection * 3 + tries


# ------------------------------------------------------------- choosers
# One section (one draw: no node action), one Sample on species 1 (particles
# 3 and 4 of 5), free particles, L = 2 in 2D.  After the chooser's draws every
# uniform is 1e-6: Box-Muller displacements stay small and every level
# accepts, so the draws after the chooser are 9 (levels 2, 1, 0: 0 + 2 + 4
# Gaussian uniforms and 3 accept draws).

def _chooser_run(chooser, prefix):
    d = capi.make_desc(2, 8, 2, 0.1, [(3, 1.0, 0.0), (2, 1.0, 0.0)],
                       samples=[(1, 1, chooser)], nsection=1)
    ch = ol.Chain(d)
    beads = capi.initial_beads(d, 3)
    ch.set_beads(beads)
    ch.rng_inject(np.array([0.3] + list(prefix) + [1e-6] * 64))
    ch.run_sweeps(1)
    moved = [p for p in range(d.npart) if not np.array_equal(ch.get_beads()[:, p], beads[:, p])]
    return ch.stats(), moved


def test_walking_chooser_draws_particle_then_veto():
    # i0 = (int)(2 * 0.75) = 1 -> particle 3 + 1; the veto uniform is below 1
    st, moved = _chooser_run(capi.PQ_CHOOSER_WALKING, [0.75, 0.999999])
    assert moved == [4]
    assert st.draws == 1 + 2 + 9 and st.tries[2] == 1


def test_walking_chooser_vetoes_only_at_one():
    # a veto uniform of exactly 1.0 (rand() == RAND_MAX) turns the move down
    st, moved = _chooser_run(capi.PQ_CHOOSER_WALKING, [0.75, 1.0])
    assert moved == [] and sum(st.tries) == 0 and st.draws == 3


def test_species_chooser_offsets_species():
    # (int)(2 * 0.25 + 3) = 3: the first particle of species 1
    st, moved = _chooser_run(capi.PQ_CHOOSER_SPECIES, [0.25])
    assert moved == [3] and st.draws == 1 + 1 + 9


def test_all_chooser_redraws_at_one():
    # (int)(5 * 1.0) == 5 is drawn again; (int)(5 * 0.5) = 2
    st, moved = _chooser_run(capi.PQ_CHOOSER_ALL, [1.0, 0.5])
    assert moved == [2] and st.draws == 1 + 2 + 9


def test_all_random_chooser_spends_one_permutation_draw():
    st, moved = _chooser_run(capi.PQ_CHOOSER_ALL_RANDOM, [0.7, 0.1])
    assert moved == [0] and st.draws == 1 + 2 + 9


def test_oracle_rejects_more_moving_particles_than_the_pool():
    # the reference's choosers would redraw forever; the oracle refuses
    d = capi.make_desc(2, 8, 2, 0.1, [(3, 1.0, 0.0)], samples=[(0, 4)])
    with pytest.raises(ValueError):
        ol.Chain(d)
    d = capi.make_desc(2, 8, 2, 0.1, [(3, 1.0, 0.0)], samples=[(0, 9)])
    with pytest.raises(ValueError):
        ol.Chain(d)
    ol.Chain(capi.make_desc(2, 8, 2, 0.1, [(3, 1.0, 0.0)], samples=[(0, 3)]))


def test_sample_nrepeat_tries_and_balance():
    """<Sample nrepeat="4">: four tries on every section (the top level is
    tried 4 x nsection times per sweep) and the same SHO energy,
    E = 0.75 coth(0.25) for omega 0.5, T 1 in 3D."""
    d = capi.make_desc(3, 32, 4, 1 / 32, [(1, 1.0, -1.0)], cell=20.0, sho=[(0, 0.5)],
                       samples=[(0, 1, capi.PQ_CHOOSER_SPECIES, 4)], nsection=2)
    ch = ol.Chain(d)
    ch.set_beads(np.zeros((32, 1, 3)))
    ch.rng_philox(3, 0)
    ch.run_sweeps(50)
    assert ch.stats().tries[4] == 50 * 2 * 4
    blocks = ch.energy_blocks(400, 100)
    err = blocks.std() / math.sqrt(len(blocks))
    assert abs(blocks.mean() - 0.75 / math.tanh(0.25)) < max(4 * err, 0.08), (blocks.mean(), err)


def _sho_seed_energies(opts, L=3, nseed=8, nblock=100, blocksize=250):
    """One SHO particle (3D, omega 0.5, T 1, M 32) per seed; the mean thermal
    energy of each seed's run and the sampler counters of the last one."""
    means, st = [], None
    for seed in range(nseed):
        d = capi.make_desc(3, 32, L, 1 / 32, [(1, 1.0, -1.0)], cell=20.0, sho=[(0, 0.5)],
                           samples=[(0, 1, capi.PQ_CHOOSER_SPECIES, 1, opts)], nsection=4)
        ch = ol.Chain(d)
        ch.set_beads(np.zeros((32, 1, 3)))
        ch.rng_philox(100 + seed, 0)
        ch.run_sweeps(200)
        means.append(ch.energy_blocks(nblock, blocksize).mean())
        st = ch.stats()
    return np.array(means), st


@pytest.mark.parametrize("opts", [{"sigma_factor": 0.9}, {"delayed_rejection": 1}])
def test_sigma_factor_and_delayed_rejection_balance(opts):
    """<Sample defaultSigmaFactor newSigmaFactor delayedRejection>
    (MultiLevelSampler.cc:100-177): a scaled mover and the delayed second
    proposal at level L-1 keep detailed balance, so independent seeds give
    E = 0.75 coth(0.25) for the 3D SHO, omega 0.5, T 1.  Delayed rejection
    shows in the bead updates: each second proposal adds one bead beyond
    what the tried levels propose."""
    means, st = _sho_seed_energies(opts)
    err = means.std(ddof=1) / math.sqrt(len(means))
    exact = 0.75 / math.tanh(0.25)
    assert abs(means.mean() - exact) < max(4 * err, 0.05), (means.mean(), err)
    L = 3
    plain = sum(st.tries[l] << (L - 1 - l) for l in range(L))
    if opts.get("delayed_rejection"):
        assert st.bead_updates > plain
    else:
        assert st.bead_updates == plain


@pytest.mark.parametrize("L,M", [(1, 4), (2, 8)])
def test_delayed_rejection_conditional_bead_variance(L, M):
    """Delayed rejection with a wide first (defaultSigmaFactor 1.5) and a
    wider second proposal (newSigmaFactor 3): with L = 1 the move is
    Tierney-Mira two-stage delayed rejection of one bead, and the exact SHO
    link action (spring + SHOAction, m = omega = 1, tau = 1/M) fixes each
    bead's law given its neighbours: Gaussian about (x_{t-1} + x_{t+1}) /
    (2 cosh tau) with variance tanh(tau)/2 per coordinate.  The sample mean
    of the squared residual must match it; using the second proposal's
    factor in the return move (C to B) instead of the first's moves it by
    more than 20 standard errors at L = 1."""
    tau = 1.0 / M
    exact = math.tanh(tau) / 2
    opts = {"delayed_rejection": 1, "sigma_factor": 1.5, "new_sigma_factor": 3.0}
    means = []
    for seed in range(4):
        d = capi.make_desc(2, M, L, tau, [(1, 1.0, -1.0)], cell=50.0, sho=[(0, 1.0)],
                           samples=[(0, 1, capi.PQ_CHOOSER_SPECIES, 1, opts)], nsection=M)
        ch = ol.Chain(d)
        ch.set_beads(np.zeros((M, 1, 2)))
        ch.rng_philox(seed, 0)
        ch.run_sweeps(100)
        acc = []
        for _ in range(15000):
            ch.run_sweeps(1)
            x = ch.get_beads()[:, 0, :]
            r = x - (np.roll(x, 1, 0) + np.roll(x, -1, 0)) / (2 * math.cosh(tau))
            acc.append(np.mean(r * r))
        means.append(np.mean(acc))
    means = np.array(means)
    err = means.std(ddof=1) / math.sqrt(len(means))
    assert abs(means.mean() - exact) < max(4 * err, 0.01 * exact), (means.mean(), err, exact)


def test_sigma_factor_changes_the_proposals_only():
    """defaultSigmaFactor 1 is the reference's default: the same draws give
    the same path as no factor; another factor changes the path."""
    def path(opts):
        d = capi.make_desc(3, 32, 3, 1 / 32, [(1, 1.0, -1.0)], cell=20.0, sho=[(0, 0.5)],
                           samples=[(0, 1, capi.PQ_CHOOSER_SPECIES, 1, opts)], nsection=4)
        ch = ol.Chain(d)
        ch.set_beads(np.zeros((32, 1, 3)))
        ch.rng_philox(5, 0)
        ch.run_sweeps(20)
        return ch.get_beads()
    base = path({})
    assert np.array_equal(path({"sigma_factor": 1.0}), base)
    assert not np.array_equal(path({"sigma_factor": 0.9}), base)


def test_many_body_nodal_distance_reduces_to_one_body():
    """useManyBodyDistance with one particle in all, of the node species: the
    distance 1/sqrt(sum 1/d^2) is d itself, so the nodal action equals the
    one-body form; with a second species it differs."""
    def run(nodes_many, species):
        d = capi.make_desc(2, 32, 2, 0.1, species, sho=[(0, 1.0)],
                           nodes=[(0, capi.PQ_NODES_SHO, 1.0)], nsection=8)
        d.nodes[0].many_body_distance = nodes_many
        ch = ol.Chain(d)
        rs = np.random.default_rng(9)
        S = 5
        sec1 = rs.normal(size=(S, d.npart, 2))
        sec2 = rs.normal(size=(S, d.npart, 2))
        mov = sec1[:, [0], :] + 0.05 * rs.normal(size=(S, 1, 2))
        mov[0], mov[-1] = sec1[0, [0]], sec1[-1, [0]]
        ch.fn_initialize(sec1, sec2)
        return ch.fn_action_difference(mov, [0], 0)
    one = [(1, 1.0, 0.0)]
    assert run(1, one) == pytest.approx(run(0, one), rel=1e-13, abs=1e-15)
    two = [(1, 1.0, 0.0), (1, 1.0, 0.0)]
    assert run(1, two) != pytest.approx(run(0, two), rel=1e-6)
