"""Pins the oracle's element operators against independent textbook assembly.

Reference routines pinned here: Internal_Force (mechanical_pcg.cpp:114-191),
PreM_u (:63-112), MP_all_GAUSS_Point (:26-61), gauss_point_MP and K_G_d_g
(material_gauss_point.cpp:18-59), Fld and Precondition_M
(phase_field_pcg.cpp:20-103), fd_h (phase_field_right_hand_side.cpp:14-47),
Dshape_function (shape_function.cpp:16-75), epsilon_routine / Sigma_g
(mechanical_strain.cpp:15-61, mechanical_stress.cpp:14-27).
"""
import numpy as np
import pytest

import oracle as O
import textbook as T

DOMAIN = (0.0, 1.5, 0.0, 3.0, -1.0, 3.0)  # hx=0.5, hy=1, hz=2 with element_nb (3,3,2)
ENB = (3, 3, 2)


def fine_oracle(seed, zero_frac=0.0, small_k=1e-5):
    o = O.Oracle(0, ENB, DOMAIN, small_k=small_k, lc=1.7)
    dims = o.dims(0)
    K, G, d, fl = T.mesh(dims, seed, zero_frac=zero_frac)
    o.set(O.K, 0, K)
    o.set(O.G, 0, G)
    o.set(O.D, 0, d)
    o.set(O.FLAG, 0, fl)
    coef = T.fine_coefficients(dims, o.h(0), K, G, d, fl, small_k)
    return o, dims, coef


@pytest.mark.parametrize("seed,zero_frac", [(1, 0.0), (2, 0.15), (3, 0.3)])
def test_internal_force_fine_matches_assembled(seed, zero_frac):
    o, dims, coef = fine_oracle(seed, zero_frac)
    A = T.elastic_matrix(dims, o.h(0), coef)
    x = np.random.default_rng(seed + 100).standard_normal((3,) + dims[::-1])
    o.set(O.P, 0, x)
    o.internal_force(0, O.P, O.W)
    y = o.vec(O.W, 0).reshape(-1)
    ref = A @ x.reshape(-1)
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_internal_force_coarse_uses_element_constants():
    o = O.Oracle(1, ENB, DOMAIN, small_k=1e-5)
    dims = o.dims(0)
    rng = np.random.default_rng(7)
    Ke = rng.uniform(1, 2, dims[::-1])
    Ge = rng.uniform(0.3, 0.9, dims[::-1])
    o.set(O.KE, 0, Ke)
    o.set(O.GE, 0, Ge)
    # nodal coefficients on the coarse level must not matter
    o.set(O.K, 0, rng.uniform(5, 6, dims[::-1]))
    o.set(O.D, 0, rng.uniform(0, 1, dims[::-1]))
    A = T.elastic_matrix(dims, o.h(0), T.coarse_coefficients(dims, Ke, Ge))
    x = rng.standard_normal((3,) + dims[::-1])
    o.set(O.P, 0, x)
    o.internal_force(0, O.P, O.W)
    ref = A @ x.reshape(-1)
    assert np.max(np.abs(o.vec(O.W, 0).reshape(-1) - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_stiffness_symmetric_and_rigid_translations():
    o, dims, coef = fine_oracle(11, 0.1)
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal((2, 3) + dims[::-1])
    o.set(O.P, 0, x)
    o.internal_force(0)
    Ax = o.vec(O.W, 0).copy()
    o.set(O.P, 0, y)
    o.internal_force(0)
    Ay = o.vec(O.W, 0).copy()
    assert abs(np.sum(y * Ax) - np.sum(x * Ay)) <= 1e-12 * np.sqrt(np.sum(Ax**2) * np.sum(y**2))
    for c in range(3):
        t = np.zeros((3,) + dims[::-1])
        t[c] = 1.0
        o.set(O.P, 0, t)
        o.internal_force(0)
        assert np.max(np.abs(o.vec(O.W, 0))) <= 1e-12 * np.max(np.abs(Ax))


def test_coarse_operator_annihilates_infinitesimal_rotations():
    o = O.Oracle(1, ENB, DOMAIN)
    dims = o.dims(0)
    o.set(O.KE, 0, np.full(dims[::-1], 2.0))
    o.set(O.GE, 0, np.full(dims[::-1], 0.7))
    hx, hy, hz = o.h(0)
    zz, yy, xx = np.meshgrid(np.arange(dims[2]) * hz, np.arange(dims[1]) * hy, np.arange(dims[0]) * hx,
                             indexing="ij")
    rot = np.stack([-yy, xx, np.zeros_like(xx)])  # rotation about z
    o.set(O.P, 0, rot)
    o.internal_force(0)
    assert np.max(np.abs(o.vec(O.W, 0))) <= 1e-12


def test_energy_of_linear_field_is_volume_times_strain_energy():
    # u = E x with symmetric E and constant moduli: x.Ax = volume * eps:sigma
    o = O.Oracle(1, ENB, DOMAIN)
    dims = o.dims(0)
    Kc, Gc = 2.5, 0.8
    o.set(O.KE, 0, np.full(dims[::-1], Kc))
    o.set(O.GE, 0, np.full(dims[::-1], Gc))
    hx, hy, hz = o.h(0)
    zz, yy, xx = np.meshgrid(np.arange(dims[2]) * hz, np.arange(dims[1]) * hy, np.arange(dims[0]) * hx,
                             indexing="ij")
    E = np.array([[0.3, 0.1, -0.2], [0.1, -0.4, 0.05], [-0.2, 0.05, 0.2]])
    X = np.stack([xx, yy, zz])
    u = np.einsum("cm,mzyx->czyx", E, X)
    o.set(O.P, 0, u)
    o.internal_force(0)
    energy = np.sum(u * o.vec(O.W, 0))
    tr = np.trace(E)
    sig = (Kc - 2 * Gc / 3) * tr * np.eye(3) + 2 * Gc * E
    vol = (dims[0] - 1) * hx * (dims[1] - 1) * hy * (dims[2] - 1) * hz
    assert energy == pytest.approx(vol * np.sum(sig * E), rel=1e-12)


@pytest.mark.parametrize("seed,zero_frac", [(21, 0.0), (22, 0.2)])
def test_prem_u_is_diagonal_of_assembled(seed, zero_frac):
    o, dims, coef = fine_oracle(seed, zero_frac)
    A = T.elastic_matrix(dims, o.h(0), coef)
    o.prem_u(0, O.W)
    assert np.allclose(o.vec(O.W, 0).reshape(-1), A.diagonal(), rtol=1e-12, atol=0)


def test_zero_node_rule_switches_whole_element_to_zero():
    # one node with K == 0 zeroes K at all Gauss points of its 8 elements
    o, dims, _ = fine_oracle(31, 0.0)
    K = o.field(O.K, 0).copy()
    K[1, 1, 1] = 0.0
    o.set(O.K, 0, K)
    o.set(O.FLAG, 0, np.ones(dims[::-1]))  # flag 1: K undegraded, G degraded
    coef = T.fine_coefficients(dims, o.h(0), K, o.field(O.G, 0), o.field(O.D, 0), o.field(O.FLAG, 0), 1e-5)
    for e in [(0, 0, 0), (1, 1, 1), (0, 1, 0)]:
        assert np.all(coef[e][0] == 0.0)
    A = T.elastic_matrix(dims, o.h(0), coef)
    x = np.random.default_rng(3).standard_normal((3,) + dims[::-1])
    o.set(O.P, 0, x)
    o.internal_force(0)
    ref = A @ x.reshape(-1)
    assert np.max(np.abs(o.vec(O.W, 0).reshape(-1) - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_degradation_closed_form_through_element_mean():
    # uniform K, G, d: MP_element gives K_e = g(d) K (flag 0), K (flag 1, 2)
    o = O.Oracle(0, (2, 2, 2), (0, 2, 0, 2, 0, 2), small_k=1e-3)
    dims = o.dims(0)
    for dval in (-0.3, 0.0, 0.4, 1.0, 1.7):
        o.set(O.K, 0, np.full(dims[::-1], 3.0))
        o.set(O.G, 0, np.full(dims[::-1], 1.2))
        o.set(O.D, 0, np.full(dims[::-1], dval))
        o.set(O.FLAG, 0, np.array([0.0, 1.0, 2.0])[np.arange(27) % 3].reshape(dims[::-1]))
        o.mp_all_element()
        dcl = min(max(dval, 0.0), 1.0)
        g = (1 - 1e-3) * (1 - dcl) ** 2 + 1e-3
        if dval == 0.0:  # zero-node rule on d: d_g = 0 -> g = 1
            g = 1.0
        fl = o.field(O.FLAG, 0)
        Ke, Ge = o.field(O.KE, 0), o.field(O.GE, 0)
        for (k, j, i) in T.all_combos(2):
            expK = 3.0 * g if fl[k, j, i] == 0 else 3.0
            expG = 1.2 if fl[k, j, i] == 2 else 1.2 * g
            assert Ke[k, j, i] == pytest.approx(expK, rel=1e-14)
            assert Ge[k, j, i] == pytest.approx(expG, rel=1e-14)


@pytest.mark.parametrize("seed", [41, 42])
def test_fld_and_precondition_m_match_assembled(seed):
    o = O.Oracle(0, ENB, DOMAIN, lc=1.3)
    dims = o.dims(0)
    rng = np.random.default_rng(seed)
    Hn = rng.uniform(0, 5, dims[::-1])
    gc = rng.uniform(0.5, 2.0, dims[::-1])
    gc[rng.random(dims[::-1]) < 0.1] = 0.0
    o.set(O.H, 0, Hn)
    o.set(O.GC, 0, gc)
    A = T.phase_matrix(dims, o.h(0), Hn, gc, 1.3)
    x = rng.standard_normal(dims[::-1])
    o.set(O.PD, 0, x)
    o.fld(0, O.PD, O.WD)
    ref = A @ x.reshape(-1)
    assert np.max(np.abs(o.field(O.WD, 0).reshape(-1) - ref)) <= 1e-12 * np.max(np.abs(ref))
    o.precondition_m(0, O.WD)
    assert np.allclose(o.field(O.WD, 0).reshape(-1), A.diagonal(), rtol=1e-12, atol=0)


def test_fld_constant_field_is_reaction_mass_row_sum():
    # with H = 0 and constant gc, A.1 = (gc/lc) * lumped mass (gradient term vanishes)
    o = O.Oracle(0, ENB, DOMAIN, lc=2.0)
    dims = o.dims(0)
    o.set(O.GC, 0, np.full(dims[::-1], 3.0))
    o.set(O.PD, 0, np.ones(dims[::-1]))
    o.fld(0)
    hx, hy, hz = o.h(0)
    count = np.zeros(dims[::-1])
    for (i, j, k) in T.elements(dims):
        count[k:k + 2, j:j + 2, i:i + 2] += 1
    assert np.allclose(o.field(O.WD, 0), 3.0 / 2.0 * count * hx * hy * hz / 8.0, rtol=1e-13)


def test_fd_h_closed_form():
    o = O.Oracle(0, ENB, DOMAIN)
    dims = o.dims(0)
    Hn = np.random.default_rng(9).uniform(0, 3, dims[::-1])
    o.set(O.H, 0, Hn)
    o.fd_h(0)
    hx, hy, hz = o.h(0)
    exp = np.zeros(dims[::-1])
    for (i, j, k) in T.elements(dims):
        exp[k:k + 2, j:j + 2, i:i + 2] += 2.0 * Hn[k, j, i] * hx * hy * hz / 8.0
    assert np.allclose(o.field(O.FD, 0), exp, rtol=1e-13)


def test_shape_functions_closed_form():
    N, DN = O.shape_functions(0.5, 1.0, 2.0)
    Nt, dNt = T.hex8_tables((0.5, 1.0, 2.0))
    # reference node order (shape_function.cpp:17-19) -> lexicographic order
    ref_nodes = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    perm = [T.LOCAL.index(n) for n in ref_nodes]
    ref_gps = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    gperm = [T.GPTS.index(g) for g in ref_gps]
    assert np.allclose(N, Nt[np.ix_(gperm, perm)], rtol=0, atol=1e-15)
    for c in range(3):
        assert np.allclose(DN[c], dNt[np.ix_(gperm, perm)][:, :, c], rtol=0, atol=1e-15)
    assert np.allclose(N.sum(axis=1), 1.0, atol=1e-15)
    assert np.allclose(DN.sum(axis=2), 0.0, atol=1e-15)


def test_reference_pointwise_build_agrees_with_oracle():
    """The reference's own Dshape_function / epsilon_routine / Sigma_g,
    compiled from its sources into oracle/_ref, against the oracle and the
    closed-form patch test."""
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes
    dp = ctypes.POINTER(ctypes.c_double)
    N = np.zeros(64)
    DN = np.zeros(192)
    R.ref_dshape_function(0.5, 1.0, 2.0, N.ctypes.data_as(dp), DN.ctypes.data_as(dp))
    No, DNo = O.shape_functions(0.5, 1.0, 2.0)
    assert np.array_equal(N.reshape(8, 8), No)
    assert np.array_equal(DN.reshape(3, 8, 8), DNo)
    # patch test: u = E x on one element
    E = np.array([[0.3, 0.1, -0.2], [0.05, -0.4, 0.05], [-0.1, 0.15, 0.2]])
    h = np.array([0.5, 1.0, 2.0])
    nodes = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    u24 = np.zeros(24)
    for a, n in enumerate(nodes):
        x = np.array(n) * h
        u24[[0 * 8 + a, 8 + a, 16 + a]] = E @ x
    eps = np.zeros(72)
    sig = np.zeros(72)
    tr = np.zeros(8)
    dev = np.zeros(8)
    R.ref_element_strain_stress(u24.ctypes.data_as(dp), 0.5, 1.0, 2.0, 2.0, 0.7, eps.ctypes.data_as(dp),
                                sig.ctypes.data_as(dp), tr.ctypes.data_as(dp), dev.ctypes.data_as(dp))
    S = 0.5 * (E + E.T)
    for g in range(8):
        assert np.allclose(eps[g * 9:(g + 1) * 9].reshape(3, 3), S, atol=1e-14)
        sref = 2.0 * np.trace(S) * np.eye(3) + 2 * 0.7 * (S - np.trace(S) / 3 * np.eye(3))
        assert np.allclose(sig[g * 9:(g + 1) * 9].reshape(3, 3), sref, atol=1e-14)


@pytest.mark.parametrize("level", [1, 0], ids=["fine", "coarse"])
def test_reference_loop_form_gives_the_same_operators(level):
    """The reference-cost loop form that times the CPU baseline
    (re-evaluated Gauss-point coefficients, critical-section scatter) applies
    the same operator and diagonal, on the finest level (nodal moduli,
    damage, split flags) and on a coarse one (element constants): equal up
    to the scatter order's rounding."""
    o = O.Oracle(1, ENB, DOMAIN, small_k=1e-5, lc=1.7)
    K, G, d, fl = T.mesh(o.dims(1), 7, zero_frac=0.1)
    for f, a in ((O.K, K), (O.G, G), (O.D, d), (O.FLAG, fl)):
        o.set(f, 1, a)
    for f in (O.K, O.G):
        o.coarse_vr_nodal(f)
    o.mech_setup()
    dims = o.dims(level)
    x = np.random.default_rng(8).standard_normal((3,) + dims[::-1])
    o.set(O.P, level, x)
    out = {}
    O.set_threads(4)
    try:
        for ref in (False, True):
            O.set_ref_loops(ref)
            o.internal_force(level, O.P, O.W)
            y = o.vec(O.W, level).copy()
            o.prem_u(level, O.W)
            out[ref] = (y, o.vec(O.W, level).copy())
    finally:
        O.set_ref_loops(False)
        O.set_threads(0)
    for a, b in zip(out[False], out[True]):
        assert np.max(np.abs(a)) > 0
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a))
