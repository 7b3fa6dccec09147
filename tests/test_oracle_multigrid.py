"""Pins the oracle's transfers, coarse coefficients, smoothers, cycles and
solves against independent matrix formulations.

Reference routines pinned: solution_injection, coarsen_f_standard_routine,
refine_standard_routine (multigrid_intergrid.cpp:25-520); coarse_voigt,
coarse_reuss, VR_all_level, Voigt_all_level (material_coarse_element.cpp:20-118);
voight, reuss, coarse_VR (image_input.cpp:144-273); pcg_u
(mechanical_pcg.cpp:223-330); pcg_d (phase_field_pcg.cpp:133-214); cycle_u
(mechanical_multigrid.cpp:18-54); cycle_d (phase_field_multigrid.cpp:18-44);
u_entire (mechanical_displacement.cpp:25-115); phase_field (phase_field.cpp:15-57).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
import textbook as T

DOM = (0.0, 2.0, 0.0, 2.0, -1.0, 1.0)


def vflat(a):
    return np.asarray(a).reshape(-1)


def make(maxlevel=1, enb=(2, 2, 2), dom=DOM, **kw):
    return O.Oracle(maxlevel, enb, dom, **kw)


def test_prolongation_is_trilinear_and_restriction_is_its_transpose():
    o = make(1, (2, 3, 2), (0, 2, 0, 3, 0, 2))
    dc, df = o.dims(0), o.dims(1)
    P = T.prolongation(dc)
    rng = np.random.default_rng(1)
    uc = rng.standard_normal(dc[::-1])
    o.set(O.P, 0, np.stack([uc] * 3))
    o.set(O.UOLD, 0, np.zeros((3,) + dc[::-1]))
    o.set(O.W, 1, np.zeros((3,) + df[::-1]))
    o.prolong(1, O.W, O.P, O.UOLD, comp=0)
    assert np.allclose(vflat(o.field(O.W, 1, 0)), P @ vflat(uc), rtol=1e-15, atol=1e-15)
    rf = rng.standard_normal(df[::-1])
    o.set(O.W, 1, np.stack([rf] * 3))
    o.set(O.P, 0, np.zeros((3,) + dc[::-1]))
    o.restrict(1, O.W, O.P, O.FU, comp=1)
    assert np.allclose(vflat(o.field(O.FU, 0, 1)), P.T @ vflat(rf), rtol=1e-14, atol=1e-14)
    # adjoint identity <R r, v> = <r, P v>
    lhs = vflat(o.field(O.FU, 0, 1)) @ vflat(uc)
    assert lhs == pytest.approx(vflat(rf) @ (P @ vflat(uc)), rel=1e-13)
    o.set(O.W, 1, np.stack([rf] * 3))
    o.injection(0, O.W, comp=2)
    assert np.array_equal(vflat(o.field(O.W, 0, 2)), T.injection(dc) @ vflat(rf))


def test_prolongation_reproduces_linear_fields():
    o = make(1, (2, 2, 2))
    dc, df = o.dims(0), o.dims(1)
    zc, yc, xc = np.meshgrid(*[np.arange(n) * 2.0 for n in dc[::-1]], indexing="ij")
    zf, yf, xf = np.meshgrid(*[np.arange(n) * 1.0 for n in df[::-1]], indexing="ij")
    f = lambda x, y, z: 0.3 * x - 1.2 * y + 0.7 * z + 0.5
    o.set(O.P, 0, np.stack([f(xc, yc, zc)] * 3))
    o.set(O.UOLD, 0, np.zeros((3,) + dc[::-1]))
    o.set(O.W, 1, np.zeros((3,) + df[::-1]))
    o.prolong(1, O.W, O.P, O.UOLD)
    assert np.allclose(o.field(O.W, 1, 0), f(xf, yf, zf), atol=1e-14)


def test_element_vr_and_voigt_closed_form():
    o = make(2, (1, 2, 1), (0, 1, 0, 2, 0, 1))
    rng = np.random.default_rng(3)
    df = o.dims(2)
    Hf = rng.uniform(0.5, 4.0, df[::-1])
    o.set(O.KE, 2, Hf)
    o.vr_all_level(O.KE)
    o.set(O.H, 2, Hf)
    o.voigt_all_level(O.H)
    d1 = o.dims(1)
    blocks = Hf[:-1, :-1, :-1].reshape(d1[2] - 1, 2, d1[1] - 1, 2, d1[0] - 1, 2)
    V = blocks.mean(axis=(1, 3, 5))
    Rm = 1.0 / (1.0 / blocks).mean(axis=(1, 3, 5))
    exp = 0.5 * (0.5 * (V + Rm) + np.sqrt(V * Rm))
    assert np.allclose(o.field(O.KE, 1)[:-1, :-1, :-1], exp, rtol=1e-14)
    assert np.all(o.field(O.KE, 1)[-1] == 0) and np.all(o.field(O.KE, 1)[:, -1] == 0)
    assert np.allclose(o.field(O.H, 1)[:-1, :-1, :-1], V, rtol=1e-14)
    assert np.all(V >= Rm * (1 - 1e-15))
    # uniform fields are fixed points of both means
    o.set(O.KE, 2, np.full(df[::-1], 2.5))
    o.vr_all_level(O.KE)
    assert np.allclose(o.field(O.KE, 0)[:-1, :-1, :-1], 2.5, rtol=1e-15)


def test_nodal_coarse_vr_closed_form():
    o = make(1, (2, 3, 2), (0, 2, 0, 3, 0, 2))
    rng = np.random.default_rng(4)
    df, dc = o.dims(1), o.dims(0)
    M = rng.uniform(1.0, 9.0, df[::-1])
    o.set(O.GC, 1, M)
    o.coarse_vr_nodal(O.GC)
    pad = np.pad(M, 1, mode="edge")
    V = np.zeros(dc[::-1])
    Rs = np.zeros(dc[::-1])
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                win = pad[dz:dz + df[2]:2, dy:dy + df[1]:2, dx:dx + df[0]:2]
                V += win / 27.0
                Rs += 1.0 / win
    R = 27.0 / Rs
    assert np.allclose(o.field(O.GC, 0), 0.5 * (0.5 * (V + R) + np.sqrt(V * R)), rtol=1e-14)


def rotation_mode(o, l):
    nx, ny, nz = o.dims(l)
    hx, hy, hz = o.h(l)
    zz, yy, xx = np.meshgrid(DOM[4] + np.arange(nz) * hz, DOM[2] + np.arange(ny) * hy,
                             DOM[0] + np.arange(nx) * hx, indexing="ij")
    return np.stack([-(yy - DOM[2]), xx - DOM[1], np.zeros_like(xx)])


def fine_mech(seed, maxlevel=1, enb=(2, 2, 2), dom=DOM, **kw):
    o = make(maxlevel, enb, dom, **kw)
    L = maxlevel
    dims = o.dims(L)
    K, G, d, fl = T.mesh(dims, seed, zero_frac=0.0, dscale=0.5)
    o.set(O.K, L, K)
    o.set(O.G, L, G)
    o.set(O.D, L, d)
    o.set(O.FLAG, L, fl)
    return o, dims, K, G, d, fl


def bc_info(o, l, Ut):
    dims = o.dims(l)
    h = o.h(l)
    vals = T.dirichlet(dims, h, (DOM[0], DOM[2], DOM[4]), (DOM[1], DOM[3], DOM[5]), Ut)
    n = 3 * int(np.prod(dims))
    fixed = np.array(sorted(vals))
    free = np.setdiff1d(np.arange(n), fixed)
    return vals, fixed, free


@pytest.mark.parametrize("nb", [1, 3, 7])
def test_pcg_u_matches_textbook_jacobi_pcg(nb):
    o, dims, K, G, d, fl = fine_mech(5)
    L = 1
    A = T.elastic_matrix(dims, o.h(L), T.fine_coefficients(dims, o.h(L), K, G, d, fl, 1e-5))
    rng = np.random.default_rng(6)
    u0 = rng.standard_normal((3,) + dims[::-1]) * 0.1
    f = rng.standard_normal((3,) + dims[::-1])
    o.set(O.U, L, u0)
    o.set(O.FU, L, f)
    Ut = 0.37
    o.pcg_u(L, nb, Ut, t=1, f_norm=1.0)
    vals, fixed, free = bc_info(o, L, Ut)
    assert len(fixed) == 2 * 25 + 2  # bottom and top faces + 2 pins
    x0 = vflat(u0).copy()
    for q, v in vals.items():
        x0[q] = v
    x, r = T.jacobi_pcg(A, vflat(f), x0, nb, free)
    assert np.allclose(vflat(o.vec(O.U, L)), x, rtol=0, atol=1e-12 * np.max(np.abs(x)))
    assert np.allclose(vflat(o.vec(O.RU, L)), r, rtol=0, atol=1e-12 * np.max(np.abs(r)))


def test_pcg_u_at_t0_sets_f_norm_from_dirichlet_reactions():
    o, dims, K, G, d, fl = fine_mech(8)
    L = 1
    A = T.elastic_matrix(dims, o.h(L), T.fine_coefficients(dims, o.h(L), K, G, d, fl, 1e-5))
    o.set(O.U, L, np.zeros((3,) + dims[::-1]))
    o.set(O.FU, L, np.zeros((3,) + dims[::-1]))
    fn = o.pcg_u(L, 2, 1.0, t=0, f_norm=0.0)
    vals, fixed, free = bc_info(o, L, 1.0)
    x0 = np.zeros(3 * int(np.prod(dims)))
    for q, v in vals.items():
        x0[q] = v
    react = (A @ x0)[fixed]
    assert fn == pytest.approx(np.sum(react**2), rel=1e-12)
    assert np.allclose(vflat(o.vec(O.FU, L))[fixed], react, rtol=1e-12)


@pytest.mark.parametrize("nb", [1, 4])
def test_pcg_d_matches_textbook(nb):
    o = make(0, (3, 2, 2), (0, 3, 0, 2, 0, 2), lc=1.5)
    dims = o.dims(0)
    rng = np.random.default_rng(12)
    Hn = rng.uniform(0, 2, dims[::-1])
    gc = rng.uniform(0.5, 2, dims[::-1])
    d0 = rng.uniform(0, 0.3, dims[::-1])
    fd = rng.standard_normal(dims[::-1])
    for f, a in ((O.H, Hn), (O.GC, gc), (O.D, d0), (O.FD, fd)):
        o.set(f, 0, a)
    fdn = o.pcg_d(0, nb)
    assert fdn == pytest.approx(np.sum(fd**2), rel=1e-14)
    A = T.phase_matrix(dims, o.h(0), Hn, gc, 1.5)
    x, r = T.jacobi_pcg(A, vflat(fd), vflat(d0).copy(), nb)
    assert np.allclose(vflat(o.field(O.D, 0)), x, atol=1e-12 * np.max(np.abs(x)))
    assert np.allclose(vflat(o.field(O.RD, 0)), r, atol=1e-12 * np.max(np.abs(r)))


def textbook_cycle_u(o, Ut, nu, mats, bcs, u, f, l, hook=None):
    """FAS gamma=1 cycle in matrix form (injection, P, P^T, Jacobi-PCG).
    hook(l, x) sees the iterate (boundary values set) before each smoother."""
    nu0, nu1, nu2 = nu
    A = mats[l]
    vals, free = bcs[l]
    def setbc(x):
        for q, v in vals.items():
            x[q] = v
    if l == 0:
        setbc(u[0])
        if hook:
            hook(0, u[0])
        u[0], r = T.jacobi_pcg(A, f[0], u[0], nu0, free)
        return r
    setbc(u[l])
    if hook:
        hook(l, u[l])
    u[l], r = T.jacobi_pcg(A, f[l], u[l], nu1, free)
    dc = o.dims(l - 1)
    Inj = T.injection(dc)
    P = T.prolongation(dc)
    uc = np.concatenate([Inj @ u[l].reshape(3, -1)[c] for c in range(3)])
    setbc_c = bcs[l - 1][0]
    for q, v in setbc_c.items():
        uc[q] = v
    uold = uc.copy()
    u[l - 1] = uc
    rr = r.reshape(3, -1)
    f[l - 1] = mats[l - 1] @ uc + np.concatenate([P.T @ rr[c] for c in range(3)])
    textbook_cycle_u(o, Ut, nu, mats, bcs, u, f, l - 1, hook)
    corr = (u[l - 1] - uold).reshape(3, -1)
    u[l] = u[l] + np.concatenate([P @ corr[c] for c in range(3)])
    setbc(u[l])
    if hook:
        hook(l, u[l])
    u[l], r = T.jacobi_pcg(A, f[l], u[l], nu2, free)
    return r


def test_cycle_u_matches_textbook_fas_cycle():
    o, dims, K, G, d, fl = fine_mech(21, maxlevel=2, enb=(2, 2, 2))
    L = 2
    rng = np.random.default_rng(22)
    u0 = rng.standard_normal((3,) + dims[::-1]) * 0.05
    o.set(O.U, L, u0)
    o.mech_setup()  # split flag from u0, element means, coarse VR
    fl = o.field(O.FLAG, L).copy()
    mats, bcs = {}, {}
    Ut = 0.25
    for l in range(L + 1):
        dl = o.dims(l)
        if l == L:
            coef = T.fine_coefficients(dl, o.h(l), K, G, d, fl, 1e-5)
        else:
            coef = T.coarse_coefficients(dl, o.field(O.KE, l), o.field(O.GE, l))
        mats[l] = T.elastic_matrix(dl, o.h(l), coef)
        vals, fixed, free = bc_info(o, l, Ut)
        bcs[l] = (vals, free)
    f = {L: np.zeros(3 * int(np.prod(dims)))}
    u = {L: vflat(u0).copy()}
    o.set(O.FU, L, np.zeros((3,) + dims[::-1]))
    r_tb = textbook_cycle_u(o, Ut, (50, 4, 2), mats, bcs, u, f, L)
    o.cycle_u(L, Ut, t=1, f_norm=1.0)
    scale = np.max(np.abs(u[L]))
    assert np.max(np.abs(vflat(o.vec(O.U, L)) - u[L])) <= 1e-10 * scale
    assert np.max(np.abs(vflat(o.vec(O.RU, L)) - r_tb)) <= 1e-9 * np.max(np.abs(r_tb))


def nodal_vr_levels(fine, o, L):
    """Coarse nodal moduli of every level by the 27-point Voigt/Reuss window
    with edge clamping, level by level from the finest (numpy closed form)."""
    out = {L: np.asarray(fine, dtype=np.float64)}
    for l in range(L, 0, -1):
        M = out[l]
        dc = o.dims(l - 1)
        pad = np.pad(M, 1, mode="edge")
        V = np.zeros(dc[::-1])
        Rs = np.zeros(dc[::-1])
        for dz in range(3):
            for dy in range(3):
                for dx in range(3):
                    win = pad[dz:dz + 2 * dc[2] - 1:2, dy:dy + 2 * dc[1] - 1:2, dx:dx + 2 * dc[0] - 1:2]
                    V += win / 27.0
                    Rs += 1.0 / win
        R = 27.0 / Rs
        out[l - 1] = 0.5 * (0.5 * (V + R) + np.sqrt(V * R))
    return out


def element_voigt_levels(fine, o, L):
    """Coarse element fields as the mean of the 8 child elements (element
    fields live in node-sized grids whose last layer is unused)."""
    out = {L: np.asarray(fine, dtype=np.float64)}
    for l in range(L, 0, -1):
        dc = o.dims(l - 1)
        E = out[l][:-1, :-1, :-1]
        blocks = E.reshape(dc[2] - 1, 2, dc[1] - 1, 2, dc[0] - 1, 2)
        c = np.zeros(dc[::-1])
        c[:-1, :-1, :-1] = blocks.mean(axis=(1, 3, 5))
        out[l - 1] = c
    return out


def textbook_cycle_d(o, nu, gamma, mats, d, f, l, post_calls=None):
    """FAS cycle of the phase field in matrix form: nu1 Jacobi-PCG iterations,
    injection, coarse right-hand side A_c d_c + P^T r, gamma coarse visits,
    trilinear correction, then nu2 restarted Jacobi-PCG calls of nu2
    iterations each.  post_calls overrides the number of post-smoother calls
    (to show that the test tells the variants apart)."""
    nu0, nu1, nu2 = nu
    A = mats[l]
    if l == 0:
        d[0], r = T.jacobi_pcg(A, f[0], d[0], nu0)
        return r
    d[l], r = T.jacobi_pcg(A, f[l], d[l], nu1)
    dc = o.dims(l - 1)
    P = T.prolongation(dc)
    d[l - 1] = T.injection(dc) @ d[l]
    dold = d[l - 1].copy()
    f[l - 1] = mats[l - 1] @ d[l - 1] + P.T @ r
    for _ in range(gamma):
        textbook_cycle_d(o, nu, gamma, mats, d, f, l - 1, post_calls)
    d[l] = d[l] + P @ (d[l - 1] - dold)
    for _ in range(nu2 if post_calls is None else post_calls):
        d[l], r = T.jacobi_pcg(A, f[l], d[l], nu2)
    return r


def phase_problem(nu, gamma, seed, L=2, enb=(2, 2, 2), dom=DOM, lc=1.5):
    o = make(L, enb, dom, lc=lc, mg_d=(nu[0], nu[1], nu[2], gamma, 12))
    dims = o.dims(L)
    rng = np.random.default_rng(seed)
    Hf = np.zeros(dims[::-1])
    Hf[:-1, :-1, :-1] = rng.uniform(0, 3, (dims[2] - 1, dims[1] - 1, dims[0] - 1))
    gcf = rng.uniform(0.5, 2, dims[::-1])
    o.set(O.H, L, Hf)
    o.voigt_all_level(O.H)
    o.set(O.GC, L, gcf)
    o.coarse_vr_nodal(O.GC)
    Hl = element_voigt_levels(Hf, o, L)
    gcl = nodal_vr_levels(gcf, o, L)
    mats = {}
    for l in range(L + 1):
        # the oracle's coarse coefficients are the closed forms
        assert np.allclose(o.field(O.H, l), Hl[l], rtol=1e-14, atol=0)
        assert np.allclose(o.field(O.GC, l), gcl[l], rtol=1e-14, atol=0)
        mats[l] = T.phase_matrix(o.dims(l), o.h(l), Hl[l], gcl[l], lc)
    d0 = rng.uniform(0, 0.4, dims[::-1])
    fd = rng.standard_normal(dims[::-1])
    o.set(O.D, L, d0)
    o.set(O.FD, L, fd)
    return o, mats, vflat(d0).copy(), vflat(fd).copy()


@pytest.mark.parametrize("nu,gamma", [((10, 2, 1), 1), ((10, 2, 3), 1), ((6, 1, 2), 2)],
                         ids=["as-shipped", "nu2=3", "W-cycle"])
def test_cycle_d_matches_textbook_fas_cycle(nu, gamma):
    """cycle_d (phase_field_multigrid.cpp:18-44) with coarsen_d, coarsen_f_d
    and refine_d (multigrid_intergrid.cpp:92-103, 345-371, 523-533) against
    the matrix-form FAS cycle on 3 levels: the coarse operators use the
    Voigt mean of H and the Voigt/Reuss nodal gc, and the post-smoother is
    nu2 calls of nu2 iterations."""
    L = 2
    o, mats, d0, fd = phase_problem(nu, gamma, 51)
    d = {L: d0.copy()}
    f = {L: fd.copy()}
    r_tb = textbook_cycle_d(o, nu, gamma, mats, d, f, L)
    fdn = o.cycle_d(L, 0.0)
    got = vflat(o.field(O.D, L))
    scale = np.max(np.abs(d[L]))
    err = np.max(np.abs(got - d[L])) / scale
    assert err <= 1e-12
    assert np.max(np.abs(vflat(o.field(O.RD, L)) - r_tb)) <= 1e-10 * np.max(np.abs(r_tb))
    # the last pcg_d call (finest post-smoother) sets fd_norm = |fd|^2
    assert fdn == pytest.approx(float(fd @ fd), rel=1e-14)
    # the coarse right-hand side is the FAS one
    assert np.allclose(vflat(o.field(O.FD, L - 1)), f[L - 1], rtol=0, atol=1e-12 * np.max(np.abs(f[L - 1])))
    if nu[2] > 1:
        # one post-smoother call instead of nu2 calls is 1e3 x the bound away
        d1 = {L: d0.copy()}
        f1 = {L: fd.copy()}
        textbook_cycle_d(o, nu, gamma, mats, d1, f1, L, post_calls=1)
        assert np.max(np.abs(got - d1[L])) / scale > 1e3 * 1e-12


def textbook_u_entire(o, L, Ut, mats, bcs, u0, f0, tol, ncy, nu, f_norm, t):
    """u_entire's loop (mechanical_displacement.cpp:72-108) over the matrix
    form: initial masked residual; at t >= 1 with f_norm != 0 the relative
    residual is r.r/f_norm, else 1; up to ncy cycles while it exceeds tol,
    each followed by the masked residual of the finest level.  At t == 0
    every finest-level smoother call stores the reactions A x of its
    starting iterate on the Dirichlet DOFs of f and sets f_norm to their
    squared norm, so the last call's value is the one the test uses."""
    vals, free = bcs[L]
    u = {L: u0.copy()}
    f = {L: f0.copy()}
    A = mats[L]
    fixed = np.setdiff1d(np.arange(len(u0)), free)
    r = f0 - A @ u0
    mask = np.zeros_like(r)
    mask[free] = 1.0
    r_norm = float(np.sum((r * mask) ** 2))
    hist = [r_norm]
    state = {"f_norm": f_norm}

    def hook(l, x):
        if t == 0 and l == L:
            react = (A @ x)[fixed]
            f[L][fixed] = react
            state["f_norm"] = float(np.sum(react ** 2))

    rel = r_norm / f_norm if (t >= 1 and f_norm != 0) else 1.0
    cycles = 0
    for _ in range(ncy):
        if rel > tol:
            r = textbook_cycle_u(o, Ut, nu, mats, bcs, u, f, L, hook)
            f_norm = state["f_norm"]
            cycles += 1
            r_norm = float(np.sum(r ** 2))
            rel = r_norm / f_norm
            hist.append(r_norm)
        if rel <= tol:
            break
    return u[L], f_norm, cycles, np.array(hist)


@pytest.mark.parametrize("tol,t,f_in", [(1e-6, 0, 0.0), (1e-12, 0, 0.0), (1e-12, 1, 3.7), ("edge", 0, 0.0),
                                        (1e-6, 0, 1e30)],
                         ids=["shipped-tol", "tight", "t1", "tol-edge", "t0-stale-fnorm"])
def test_u_entire_stopping_logic_matches_textbook(tol, t, f_in):
    """Cycle count, residual history and f_norm of u_entire against the
    matrix-form loop: pins the squared relative test (tol-edge puts the
    tolerance just above the 5th cycle's relative residual, so the solve
    stops there and not one cycle later), the f_norm of t == 0 (whatever
    value comes in), the t >= 1 first-residual rule and the ncy cap."""
    ncy = 8
    edge = tol == "edge"
    o, dims, K, G, d, fl = fine_mech(61, maxlevel=2, enb=(2, 2, 2), tol_u=1e-30 if edge else tol,
                                     mg_u=(50, 4, 2, 1, ncy))
    L = 2
    Ut = 1.0
    u0 = np.zeros((3,) + dims[::-1])
    if t == 1:
        u0 = np.random.default_rng(62).standard_normal((3,) + dims[::-1]) * 1e-3
    o.set(O.U, L, u0)
    o.set(O.FU, L, np.zeros((3,) + dims[::-1]))
    o.mech_setup()
    flags = o.field(O.FLAG, L).copy()
    mats, bcs = {}, {}
    for l in range(L + 1):
        dl = o.dims(l)
        coef = (T.fine_coefficients(dl, o.h(l), K, G, d, flags, 1e-5) if l == L
                else T.coarse_coefficients(dl, o.field(O.KE, l), o.field(O.GE, l)))
        mats[l] = T.elastic_matrix(dl, o.h(l), coef)
        vals, fixed, free = bc_info(o, l, Ut)
        bcs[l] = (vals, free)
    f0 = np.zeros(3 * int(np.prod(dims)))
    if edge:
        _, fn1, _, h1 = textbook_u_entire(o, L, Ut, mats, bcs, vflat(u0), f0, 1e-30, ncy, (50, 4, 2), f_in, t)
        tol = 1.25 * h1[5] / fn1
        o = make(L, (2, 2, 2), DOM, tol_u=tol, mg_u=(50, 4, 2, 1, ncy))
        for fld, a in ((O.K, K), (O.G, G), (O.D, d)):
            o.set(fld, L, a)
        o.set(O.U, L, u0)
        o.mech_setup()
    u_tb, fn_tb, cyc_tb, hist_tb = textbook_u_entire(o, L, Ut, mats, bcs, vflat(u0), f0, tol, ncy, (50, 4, 2), f_in, t)
    o.set(O.U, L, u0)
    o.set(O.FU, L, np.zeros((3,) + dims[::-1]))
    fn, cyc = o.u_entire(Ut, t, f_in)
    hist = o.history("u")
    assert cyc == cyc_tb
    assert fn == pytest.approx(fn_tb, rel=1e-12)
    assert len(hist) == len(hist_tb)
    # the two formulations round differently and the 50-iteration coarsest
    # CG amplifies that from cycle to cycle: a few 1e-6 after 8 cycles
    assert np.allclose(hist, hist_tb, rtol=1e-4, atol=1e-30)
    if edge:
        assert cyc == 5
    elif tol == 1e-6:
        assert cyc < ncy  # stopped by the test, not the cap
    else:
        assert cyc == ncy  # stopped by the cap
    m = rotation_mode(o, L).reshape(-1)
    e = vflat(o.vec(O.U, L)) - u_tb
    e -= (e @ m) / (m @ m) * m
    assert np.linalg.norm(e) <= 1e-9 * np.linalg.norm(u_tb)


def textbook_phase_field(o, L, mats, d0, fd, nu, gamma, ncy, tol_rel, tol_abs):
    """phase_field's loop (phase_field.cpp:19-56): absolute test on the
    squared residual, relative test against fd_norm (= |fd|^2 after the
    last smoother call), residual read from the smoothers' updated residual."""
    d = {L: d0.copy()}
    f = {L: fd.copy()}
    r = fd - mats[L] @ d0
    r_norm = float(r @ r)
    hist = [r_norm]
    sres = 1.0
    cycles = 0
    fd_norm = 0.0
    for _ in range(ncy):
        if r_norm > tol_abs:
            if sres > tol_rel:
                r = textbook_cycle_d(o, nu, gamma, mats, d, f, L)
                fd_norm = float(fd @ fd)
                cycles += 1
            r_norm = float(r @ r)
            sres = r_norm / fd_norm
            hist.append(r_norm)
            if sres <= tol_rel:
                break
        else:
            break
    return d[L], cycles, np.array(hist)


@pytest.mark.parametrize("tol_rel,tol_abs", [(1e-12, 1e-12), (1e-24, 1e-40), (1e-12, 1e3)],
                         ids=["shipped", "cap", "abs-stop"])
def test_phase_field_stopping_logic_matches_textbook(tol_rel, tol_abs):
    L = 2
    nu = (10, 2, 1)
    ncy = 6
    o, mats, d0, fd = phase_problem(nu, 1, 71)
    o2 = make(L, (2, 2, 2), DOM, lc=1.5, mg_d=(10, 2, 1, 1, ncy), tol_d_rel=tol_rel, tol_d_abs=tol_abs)
    for l in range(L + 1):
        o2.set(O.H, l, o.field(O.H, l))
        o2.set(O.GC, l, o.field(O.GC, l))
    o2.set(O.D, L, d0.reshape(o.dims(L)[::-1]))
    o2.set(O.FD, L, fd.reshape(o.dims(L)[::-1]))
    d_tb, cyc_tb, hist_tb = textbook_phase_field(o, L, mats, d0, fd, nu, 1, ncy, tol_rel, tol_abs)
    fdn, cyc = o2.phase_field(0, 0.0)
    hist = o2.history("d")
    assert cyc == cyc_tb
    assert len(hist) == len(hist_tb)
    assert np.allclose(hist, hist_tb, rtol=1e-6, atol=0)
    assert np.linalg.norm(vflat(o2.field(O.D, L)) - d_tb) <= 1e-11 * np.linalg.norm(d_tb)
    if tol_abs == 1e3:
        assert cyc == 0 and len(hist) == 1
    elif tol_rel == 1e-24:
        assert cyc == ncy
    else:
        assert 0 < cyc < ncy


def test_u_entire_converges_to_direct_solve():
    o, dims, K, G, d, fl = fine_mech(31, maxlevel=2, enb=(2, 2, 2), tol_u=1e-26,
                                     mg_u=(50, 4, 2, 1, 40))
    L = 2
    o.set(O.U, L, np.zeros((3,) + dims[::-1]))
    fnorm, cycles = o.u_entire(1.0, 0, 0.0)
    assert 1 <= cycles <= 40
    # at t=0 from u=0 every split flag is 0
    assert np.all(o.field(O.FLAG, L)[:-1, :-1, :-1] == 0)
    A = T.elastic_matrix(dims, o.h(L), T.fine_coefficients(dims, o.h(L), K, G, d, np.zeros(dims[::-1]), 1e-5))
    vals, fixed, free = bc_info(o, L, 1.0)
    n = A.shape[0]
    uc = np.zeros(n)
    for q, v in vals.items():
        uc[q] = v
    rhs = -(A @ uc)[free]
    got = vflat(o.vec(O.U, L))
    # The reference's pins (u_x at (x_mid, y_a), u_y at (x_b, y_mid)) leave
    # one zero-energy mode: the rotation about the vertical axis through the
    # corner (x_b, y_a).  The direct solve fixes it with a rank-one penalty
    # and solutions are compared modulo that mode.
    m = rotation_mode(o, L).reshape(-1)
    assert np.linalg.norm(A @ m) <= 1e-12 * np.linalg.norm(A.diagonal()) * np.linalg.norm(m)
    Aff = A[free][:, free].toarray()
    mf = m[free]
    Aff += np.mean(np.diag(Aff)) * np.outer(mf, mf) / (mf @ mf)
    x = uc.copy()
    x[free] = np.linalg.solve(Aff, rhs)
    e = got - x
    e -= (e @ m) / (m @ m) * m
    # the FAS cycle stagnates near a squared relative residual of 1e-19 here,
    # so the error bound follows from the true residual of the iterate
    res = np.linalg.norm((A @ got)[free])
    assert res <= 2e-9 * np.sqrt(fnorm)
    assert np.linalg.norm(e) <= 1e-7 * np.linalg.norm(x)
    # f_norm is the squared reaction norm on the Dirichlet DOFs
    assert fnorm == pytest.approx(np.sum((A @ x)[fixed] ** 2), rel=1e-8)
    # externe_force_nodal stores A u on the Dirichlet DOFs
    assert np.allclose(vflat(o.vec(O.FU, L))[fixed], (A @ got)[fixed], rtol=1e-10)
    hist = o.history("u")
    assert hist[-1] / fnorm <= 1e-17


def test_phase_field_converges_to_direct_solve():
    o = make(2, (2, 2, 2), DOM, lc=1.5, mg_d=(10, 2, 1, 1, 40), tol_d_rel=1e-24, tol_d_abs=1e-30)
    L = 2
    dims = o.dims(L)
    rng = np.random.default_rng(41)
    Hf = rng.uniform(0, 3, dims[::-1])
    gcs = [rng.uniform(0.5, 2, o.dims(l)[::-1]) for l in range(L + 1)]
    o.set(O.H, L, Hf)
    o.voigt_all_level(O.H)
    o.set(O.GC, L, gcs[L])
    o.coarse_vr_nodal(O.GC)
    o.fd_h(L)
    fdn, cycles = o.phase_field(0, 0.0)
    A = T.phase_matrix(dims, o.h(L), Hf, gcs[L], 1.5)
    x = spla.spsolve(A.tocsc(), vflat(o.field(O.FD, L)))
    assert np.linalg.norm(vflat(o.field(O.D, L)) - x) <= 1e-10 * np.linalg.norm(x)
    assert fdn == pytest.approx(np.sum(o.field(O.FD, L) ** 2), rel=1e-14)
    assert 1 <= cycles < 40


def test_cycle_from_rough_start_is_rounding_sensitive():
    """Documents a property of the reference algorithm that bounds what any
    re-implementation can match: its boundary conditions leave a rigid
    rotation mode (see test_u_entire_converges_to_direct_solve) and the
    coarsest smoother runs a fixed 50 CG iterations, far past convergence on a
    5^3 grid.  On a two-phase voxel material with damage (zeros included),
    from a rough random start, a 1e-14 perturbation of the input changes the
    cycle's result at the 1e-5 level."""
    from gpu_common import material_from_labels, spheres_labels

    def run(eps):
        o = make(1, (4, 4, 4), (0.0, 8.0, 0.0, 8.0, 0.0, 8.0), lc=2.0)
        dims = o.dims(1)
        K, G, gc = material_from_labels(spheres_labels(dims, 3, 1.5, 2.0, 13))
        for f, a in ((O.K, K), (O.G, G), (O.GC, gc)):
            o.set(f, 1, a)
        for f in (O.K, O.G, O.GC):
            o.coarse_vr_nodal(f)
        rng = np.random.default_rng(14)
        d = rng.uniform(-0.05, 0.9, dims[::-1])
        d[rng.random(dims[::-1]) < 0.1] = 0.0
        o.set(O.D, 1, d)
        u0 = np.random.default_rng(15).standard_normal((3,) + dims[::-1]) * 0.01 * (1 + eps)
        o.set(O.U, 1, u0)
        o.set(O.FU, 1, np.zeros((3,) + dims[::-1]))
        o.mech_setup()
        fn = o.cycle_u(1, 1.0, 0, 0.0)
        return fn, vflat(o.vec(O.U, 1)).copy()
    fa, ua = run(0.0)
    fb, ub = run(1e-14)
    assert np.linalg.norm(ua - ub) / np.linalg.norm(ua) > 1e-8
    assert abs(fa - fb) / fa > 1e-6


def test_solve_at_17_cubed_amplifies_a_one_ulp_load_change():
    """The physical-start counterpart of the property above, on the 17^3 and
    33^3 two-phase cases of the GPU solve-parity tests: u_entire with Ut = 1
    and with Ut = 1 + 2^-52 (an exact-arithmetic change of 1e-16).  At 17^3
    the result moves by a few 1e-10, which is the floor against which any
    re-implementation of the algorithm can be compared there; at 33^3 it
    moves by a few 1e-11."""
    from gpu_common import material_from_labels, rel_l2, spheres_labels

    def run(ml, Ut):
        dom = (0.0, 4.0 * 2 ** ml, 0.0, 4.0 * 2 ** ml, 0.0, 4.0 * 2 ** ml)
        o = make(ml, (4, 4, 4), dom, lc=2.0, small_k=1e-6, tol_u=1e-6, mg_u=(50, 4, 2, 1, 30),
                 mg_d=(10, 2, 1, 1, 30))
        dims = o.dims(ml)
        K, G, gc = material_from_labels(spheres_labels(dims, 4, 1.5, max(2.0, min(dims) / 5.0), 20))
        for f, a in ((O.K, K), (O.G, G), (O.GC, gc)):
            o.set(f, ml, a)
        for f in (O.K, O.G, O.GC):
            o.coarse_vr_nodal(f)
        o.u_entire(Ut, 0, 0.0)
        return o.vec(O.U, ml) / Ut

    ut = np.nextafter(1.0, 2.0)
    d17 = rel_l2(run(2, ut), run(2, 1.0))
    d33 = rel_l2(run(3, ut), run(3, 1.0))
    assert 1e-11 < d17 < 1e-8
    assert d33 < d17 / 3


def test_solve_at_17_cubed_moves_under_one_ulp_modulus_changes():
    """The same 17^3 case with every nodal modulus changed by at most one ulp
    (a random factor 1 + {-1, 0, 1} * 2^-52: the size of the rounding that a
    re-implementation with other summation orders and interpolation forms
    introduces).  Modulo the zero-energy rotation the five-cycle result moves
    by a median of 7e-10 over eight such changes (max 3.6e-9; at 33^3 the
    median is 1.4e-11), so a GPU-oracle distance of that size at 17^3 is the
    oracle's own rounding floor, not a defect."""
    from gpu_common import material_from_labels, spheres_labels
    from staggered import rel_l2_mod_rotation

    ml = 2
    dom = (0.0, 16.0, 0.0, 16.0, 0.0, 16.0)

    def run(pert):
        o = make(ml, (4, 4, 4), dom, lc=2.0, small_k=1e-6, tol_u=1e-6, mg_u=(50, 4, 2, 1, 30),
                 mg_d=(10, 2, 1, 1, 30))
        dims = o.dims(ml)
        K, G, gc = material_from_labels(spheres_labels(dims, 4, 1.5, max(2.0, min(dims) / 5.0), 20))
        if pert is not None:
            r = np.random.default_rng(pert)
            K = K * (1 + 2.0 ** -52 * r.choice([-1, 0, 1], K.shape))
            G = G * (1 + 2.0 ** -52 * r.choice([-1, 0, 1], G.shape))
        for f, a in ((O.K, K), (O.G, G), (O.GC, gc)):
            o.set(f, ml, a)
        for f in (O.K, O.G, O.GC):
            o.coarse_vr_nodal(f)
        o.u_entire(1.0, 0, 0.0)
        return o.vec(O.U, ml).copy(), dims

    base, dims = run(None)
    d = [rel_l2_mod_rotation(run(p)[0], base, dims, dom) for p in range(4)]
    assert 2e-10 < float(np.median(d)) < 1e-8
