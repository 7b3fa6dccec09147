"""Mixed-precision form of the mechanical solve (FP64 flexible CG around an
FP32 V-cycle) against the FP64 solution of the same discrete problem.

The reference solution is a direct solve of the independently assembled
system (tests/textbook.py: Voigt B-matrix hex8 assembly, the reference's
Dirichlet DOFs) with the zero-energy rotation mode that the reference's
boundary conditions leave fixed by one extra constraint; displacement
differences are compared with that mode projected out.  The FAS solve of the
oracle is not used as the reference here: it stagnates near a squared
relative residual of 1e-19, above what the mixed solve reaches.

Checks: the requested squared relative residual is reached (recomputed in
FP64 from u); the error against the direct solve follows the residual; the
Dirichlet reactions and f_norm match; a damaged state and a t >= 1 step
(f_norm frozen, split flags from the previous u) behave as in u_entire; two
runs are bit-identical; bad parameters are refused.  The same checks run
with the V-cycle in FP64 (precond_fp64), the like-for-like comparison of the
two preconditioner precisions."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
import textbook as T
from gpu_common import make_pair, material_from_labels, require_gpu, say

pytestmark = pytest.mark.gpu

KW = dict(lc=2.0, small_k=1e-6, tol_u=1e-16, mg_u=(50, 4, 2, 1, 30), mg_d=(10, 2, 1, 1, 30))
# maxlevel, element_nb, domain (even element counts in x and y on every level,
# so that both pins exist)
CASES = [
    (2, (4, 4, 4), (0.0, 16.0, 0.0, 16.0, 0.0, 16.0)),
    (2, (2, 4, 3), (-1.0, 7.0, 0.0, 16.0, 0.5, 12.5)),
]
IDS = ["17c", "9x17x13"]
PREC = pytest.mark.parametrize("p64", [False, True], ids=["f32", "f64"])
# preconditioner forms: FP32 (element contraction on the vector ALUs), FP64,
# FP32 with the element contraction on the matrix cores
FORMS = pytest.mark.parametrize("form", [dict(), dict(precond_fp64=True), dict(element_mfma=True),
                                         dict(coarse_iters=12)],
                                ids=["f32", "f64", "f32-mfma", "f32-coarse12"])


@pytest.fixture(scope="module")
def HA():
    return require_gpu()


def rotation_mode(dims, h, dom):
    """Rotation about the vertical axis through (x_b, y_a)."""
    nx, ny, nz = dims
    zz, yy, xx = np.meshgrid(dom[4] + np.arange(nz) * h[2], dom[2] + np.arange(ny) * h[1],
                             dom[0] + np.arange(nx) * h[0], indexing="ij")
    return np.stack([-(yy - dom[2]), xx - dom[1], np.zeros_like(xx)])


def err_modulo_rotation(u, ref, m):
    du = (np.asarray(u) - ref).ravel()
    mv = m.ravel()
    du = du - (du @ mv) / (mv @ mv) * mv
    return np.linalg.norm(du) / np.linalg.norm(ref)


def direct_solution(dims, h, dom, K, G, d, flag, Ut, small_k):
    """u with the reference's Dirichlet values solving the free rows of A u = 0
    (one u_y DOF off the rotation axis held at 0 to remove the zero mode), the
    matrix A and the Dirichlet DOF mask."""
    A = T.elastic_matrix(dims, h, T.fine_coefficients(dims, h, K, G, d, flag, small_k)).tocsr()
    vals = T.dirichlet(dims, h, (dom[0], dom[2], dom[4]), (dom[1], dom[3], dom[5]), Ut)
    n = A.shape[0]
    Nn = n // 3
    u = np.zeros(n)
    for q, v in vals.items():
        u[q] = v
    m = rotation_mode(dims, h, dom).ravel()
    cand = [q for q in range(Nn, 2 * Nn) if q not in vals]
    extra = max(cand, key=lambda q: abs(m[q]))
    free = np.array([q for q in range(n) if q not in vals and q != extra])
    rhs = -(A @ u)[free]
    u[free] = spla.spsolve(A[free][:, free].tocsc(), rhs)
    fixed = np.zeros(n, dtype=bool)
    fixed[list(vals)] = True
    return u.reshape((3,) + tuple(dims[::-1])), A, fixed


def zero_u(s, L):
    for c in range(3):
        s.fill(O.U, L, 0.0, c)
        s.fill(O.FU, L, 0.0, c)


@FORMS
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_mixed_matches_direct_solve(HA, case, form):
    ml, enb, dom = case
    say(f"[gpu] mixed vs direct solve {case}, preconditioner {form}")
    o, s, lab = make_pair(HA, ml, enb, dom, seed=40, lab_incl=4, **KW)
    K, G, _ = material_from_labels(lab)
    L = ml
    dims, h = o.dims(L), o.h(L)
    d = np.zeros(dims[::-1])
    # at t = 0 from u = 0 every split flag is 0
    ud, A, fixed = direct_solution(dims, h, dom, K, G, d, np.zeros(dims[::-1]), 1.0, KW["small_k"])
    m = rotation_mode(dims, h, dom)
    Au = A @ ud.ravel()
    f_ref = np.sum(Au[fixed] ** 2)
    errs = []
    for tol in (1e-16, 1e-22):
        zero_u(s, L)
        fs, it = s.u_entire_mixed(1.0, 0, 0.0, tol=tol, max_iter=200, **form)
        rel = s.last_stats.r_norm_final / fs
        err = err_modulo_rotation(s.get(O.U, L), ud, m)
        errs.append(err)
        say(f"  tol {tol:.0e}: {it} iterations, squared rel residual {rel:.3e}, rel L2 error {err:.3e}")
        assert rel <= tol and it < 200
        assert fs == pytest.approx(f_ref, rel=1e-6)
    # the error follows the residual: sqrt(1e-22) = 1e-11 relative residual
    assert errs[1] < 1e-9
    assert errs[1] <= errs[0]
    # Dirichlet reactions (f := A u on Dirichlet DOFs)
    fu = s.get(O.FU, L).ravel()
    assert np.linalg.norm(fu[fixed] - Au[fixed]) <= 1e-9 * np.linalg.norm(Au[fixed])


@PREC
def test_mixed_damaged_state_and_later_step(HA, p64):
    """Damage d in [-0.05, 1.05] on 30 % of the nodes (zeros elsewhere), then
    a t = 1 step: split flags from the step-0 solution, f_norm frozen."""
    ml, enb, dom = CASES[0]
    say(f"[gpu] mixed damaged state + step 1, 17^3, FP64 V-cycle {p64}")
    o, s, lab = make_pair(HA, ml, enb, dom, seed=41, lab_incl=5, **KW)
    K, G, _ = material_from_labels(lab)
    L = ml
    dims, h = o.dims(L), o.h(L)
    rng = np.random.default_rng(42)
    d = rng.uniform(-0.05, 1.05, dims[::-1]) * (rng.random(dims[::-1]) < 0.3)
    s.set(O.D, L, d)
    m = rotation_mode(dims, h, dom)
    zero_u(s, L)
    f0, it0 = s.u_entire_mixed(1.0, 0, 0.0, tol=1e-22, max_iter=200, precond_fp64=p64)
    ud0, _, _ = direct_solution(dims, h, dom, K, G, d, np.zeros(dims[::-1]), 1.0, KW["small_k"])
    e0 = err_modulo_rotation(s.get(O.U, L), ud0, m)
    f1, it1 = s.u_entire_mixed(1.7, 1, f0, tol=1e-22, max_iter=200, precond_fp64=p64)
    flags = s.get(O.FLAG, L)
    ud1, _, _ = direct_solution(dims, h, dom, K, G, d, flags, 1.7, KW["small_k"])
    e1 = err_modulo_rotation(s.get(O.U, L), ud1, m)
    say(f"  step 0: {it0} iterations, error {e0:.3e}; step 1: {it1} iterations, error {e1:.3e}, "
        f"flags {np.unique(flags[:-1, :-1, :-1])}")
    assert f1 == f0
    assert s.last_stats.r_norm_final / f0 <= 1e-22
    assert e0 < 1e-9 and e1 < 1e-9


@PREC
def test_mixed_reaches_tolerance_at_33_and_is_deterministic(HA, p64):
    ml, enb, dom = 3, (4, 4, 4), (-16.0, 16.0, -16.0, 16.0, -16.0, 16.0)
    say(f"[gpu] mixed 33^3: tolerance, iterations, determinism, FP64 V-cycle {p64}")
    _, s, _ = make_pair(HA, ml, enb, dom, seed=43, lab_incl=4, **KW)
    L = ml
    out = []
    for _ in range(2):
        zero_u(s, L)
        fs, it = s.u_entire_mixed(1.0, 0, 0.0, tol=1e-16, precond_fp64=p64)
        out.append((fs, it, s.get(O.U, L).copy(), list(s.last_stats.history[:it])))
        assert s.last_stats.r_norm_final / fs <= 1e-16 and it <= 40
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])
    assert out[0][3] == out[1][3]


@pytest.mark.parametrize("geo", [
    (2, (4, 4, 4), (0.0, 16.0, 0.0, 16.0, 0.0, 16.0)),
    (2, (2, 4, 3), (-1.0, 7.0, 0.0, 12.0, 0.5, 12.5)),
    (3, (4, 4, 4), (-16.0, 16.0, -16.0, 16.0, -16.0, 16.0)),
], ids=["17c", "aniso-9x17x13", "33c"])
def test_fp32_operator_on_vector_alus_and_matrix_cores_matches_fp64(HA, geo):
    """The preconditioner's FP32 finest-level operator, with the element
    contraction on the vector ALUs and on the matrix cores, against the FP64
    operator (itself pinned to the oracle) on a damaged state with all three
    split flags and a random vector: FP32 rounding only.  The anisotropic
    case (hx != hy) runs the scaled form, the cubic ones the folded scale."""
    ml, enb, dom = geo
    say(f"[gpu] FP32 operator (vector / matrix cores) vs FP64, {geo}")
    o, s, _ = make_pair(HA, ml, enb, dom, seed=45, lab_incl=4, **KW)
    L = ml
    dims, h = o.dims(L), o.h(L)
    rng = np.random.default_rng(46)
    shape = dims[::-1]
    s.set(O.D, L, rng.uniform(-0.05, 1.05, shape) * (rng.random(shape) < 0.4))
    s.set(O.FLAG, L, rng.integers(0, 3, shape).astype(np.float64))
    s.set(O.P, L, rng.standard_normal((3,) + shape))
    s.internal_force(L, O.P, O.W)
    y64 = s.get(O.W, L).ravel()
    s.internal_force_f32(O.P, O.W, mfma=False)
    yv = -s.get(O.W, L).ravel()
    s.internal_force_f32(O.P, O.W, mfma=True)
    ym = -s.get(O.W, L).ravel()
    fixed = np.zeros(y64.size, dtype=bool)
    fixed[list(T.dirichlet(dims, h, (dom[0], dom[2], dom[4]), (dom[1], dom[3], dom[5]), 0.0))] = True
    ev = np.linalg.norm(yv[~fixed] - y64[~fixed]) / np.linalg.norm(y64[~fixed])
    em = np.linalg.norm(ym[~fixed] - y64[~fixed]) / np.linalg.norm(y64[~fixed])
    mv = np.max(np.abs(yv[~fixed] - y64[~fixed])) / np.max(np.abs(y64[~fixed]))
    mm = np.max(np.abs(ym[~fixed] - y64[~fixed])) / np.max(np.abs(y64[~fixed]))
    say(f"  vector ALUs: rel L2 {ev:.2e}, max {mv:.2e}; matrix cores: rel L2 {em:.2e}, max {mm:.2e}")
    assert fixed.any() and not np.any(yv[fixed]) and not np.any(ym[fixed])
    assert ev < 1e-5 and em < 1e-5 and mv < 1e-5 and mm < 1e-5


def test_mixed_rejects_bad_parameters(HA):
    ml, enb, dom = CASES[0]
    _, s, _ = make_pair(HA, ml, enb, dom, seed=44, **KW)
    # cheb_degree 0 is the contrast-chosen setting; negative and > 8 are refused
    for deg in (-1, 9):
        with pytest.raises(HA.HpcrError):
            s.u_entire_mixed(1.0, 0, 0.0, cheb_degree=deg)
    with pytest.raises(HA.HpcrError):
        s.u_entire_mixed(1.0, 0, 0.0, cheb_degree=2, cheb_ratio=1.0)
    for bad in ((2, 0, 0), (0, 2, 0), (1, 1, 0), (0, 0, -1)):  # (precond_fp64, element_mfma, coarse_iters)
        mp = HA.capi.MixedParams(10, 1e-16, 2, 10, 6.0, 1.1, 10, *bad)
        fn = HA.capi.ctypes.c_double(0.0)
        st = HA.capi.Stats()
        rc = s._L.hpcr_u_entire_mixed(s._h, 1.0, 0, HA.capi.ctypes.byref(fn), HA.capi.ctypes.byref(mp),
                                      HA.capi.ctypes.byref(st))
        assert rc == HA.capi.ERR_ARG


def test_zero_load_and_zero_f_norm(HA):
    """f_norm == 0: at t == 0 with a zero load (u = 0, f = 0) the residual is
    exactly 0, so the solve is converged with no iteration and u stays 0; at
    t >= 1 with f_norm == 0 and a nonzero residual the relative residual is
    never below the tolerance, so the solve runs to its iteration cap and
    returns normally, as u_entire runs its ncy cycles
    (mechanical_displacement.cpp:84-108)."""
    ml, enb, dom = CASES[0]
    o, s, _ = make_pair(HA, ml, enb, dom, seed=3, lab_incl=2, **KW)
    L = ml
    say("[gpu] mixed solve with f_norm == 0")
    fn, it = s.u_entire_mixed(0.0, 0, 0.0, max_iter=20)
    assert fn == 0.0 and it == 0
    assert not np.any(s.get(O.U, L)) and s.last_stats.r_norm_final == 0.0
    rng = np.random.default_rng(4)
    s.set(O.U, L, rng.standard_normal((3,) + s.dims(L)[::-1]) * 1e-3)
    fn, it = s.u_entire_mixed(1.0, 1, 0.0, max_iter=7)
    assert fn == 0.0 and it == 7
    assert np.all(np.isfinite(s.get(O.U, L)))


def test_auto_chebyshev_setting_follows_modulus_contrast(HA):
    """cheb_degree 0 resolves from max K / min K: the two-phase image (contrast
    K_incl / K_matrix = 38.9e9 / 2.5e9) keeps degree 2, range 6, safety 1.1;
    with near-void nodes (E = 1e-6 E_matrix) it takes degree 4, range 12,
    safety 1.3.  An auto solve is bit-identical to the explicit call."""
    ml, enb, dom = CASES[0]
    say("[gpu] mixed smoother setting from the modulus contrast")
    o, s, lab = make_pair(HA, ml, enb, dom, seed=41, lab_incl=4, **KW)
    L = ml
    K, G, gc = material_from_labels(lab)
    kinc, kmat = 70e9 / 1.8 * 1e-6, 3e9 / 1.2 * 1e-6
    deg, ratio, up, k = s.mixed_resolve()
    assert (deg, ratio, up) == (2, 6.0, 1.1)
    assert k == pytest.approx(kinc / kmat, rel=1e-12)
    assert s.mixed_resolve(3, 8.0, 1.2) == (3, 8.0, 1.2, 0.0)

    def solve(**kw):
        zero_u(s, L)
        fs, it = s.u_entire_mixed(1.0, 0, 0.0, tol=1e-16, max_iter=200, **kw)
        return s.get(O.U, L), it, s.last_stats.r_norm_final / fs

    ua, ita, ra = solve()
    ue, ite, re = solve(cheb_degree=2, cheb_ratio=6.0, cheb_upper=1.1)
    assert ita == ite and ra <= 1e-16 and np.array_equal(ua, ue)
    # a pore: a ball of nodes at 1e-6 of the matrix modulus
    dims = o.dims(L)
    zz, yy, xx = np.meshgrid(*(np.arange(n) for n in dims[::-1]), indexing="ij")
    pore = (xx - 5) ** 2 + (yy - 11) ** 2 + (zz - 8) ** 2 <= 6
    K2, G2 = np.where(pore, K * 1e-6, K), np.where(pore, G * 1e-6, G)
    for f, a in ((O.K, K2), (O.G, G2), (O.GC, gc)):
        s.set(f, L, a)
    for f in (O.K, O.G, O.GC):
        s.coarse_vr(f)
    deg, ratio, up, k = s.mixed_resolve()
    assert (deg, ratio, up) == (4, 12.0, 1.3)
    assert k == pytest.approx(K.max() / (K[pore].min() * 1e-6), rel=1e-12)
    ua, ita, ra = solve()
    ue, ite, re = solve(cheb_degree=4, cheb_ratio=12.0, cheb_upper=1.3)
    say(f"  pore image: {ita} iterations, squared rel residual {ra:.3e}")
    assert ita == ite and ra <= 1e-16 and np.array_equal(ua, ue)
