"""Independent textbook finite-element formulations used to pin the oracle.

Nothing here reuses the oracle's loops: element matrices are built from the
standard Voigt B-matrix with 2x2x2 Gauss quadrature, assembled into scipy
sparse matrices in a node ordering of their own, and the multigrid transfer
operators are built as Kronecker products of 1-D linear interpolation.  Only
the coefficient rules that define the reference's model (the Gauss-point
interpolation with its zero-node rule, the (1-k)(1-d)^2+k degradation and the
split flag that selects it) are restated here, each from its description.
"""
import itertools

import numpy as np
import scipy.sparse as sp

GP = np.array([-1.0, 1.0]) / np.sqrt(3.0)
# local node (ix, iy, iz) in {0,1}^3, enumerated lexicographically (x fastest)
LOCAL = [(ix, iy, iz) for iz in (0, 1) for iy in (0, 1) for ix in (0, 1)]
GPTS = [(gx, gy, gz) for gz in (0, 1) for gy in (0, 1) for gx in (0, 1)]


def lin(i, xi):
    """1-D linear shape function on [-1, 1]: node 0 at -1, node 1 at +1."""
    return 0.5 * (1 - xi) if i == 0 else 0.5 * (1 + xi)


def dlin(i):
    return -0.5 if i == 0 else 0.5


def hex8_tables(h):
    """N[g, a] and dN/dx[g, a, c] at the 8 Gauss points for cell sizes h."""
    hx, hy, hz = h
    N = np.zeros((8, 8))
    dN = np.zeros((8, 8, 3))
    for g, (gx, gy, gz) in enumerate(GPTS):
        xi, eta, zeta = GP[gx], GP[gy], GP[gz]
        for a, (ix, iy, iz) in enumerate(LOCAL):
            N[g, a] = lin(ix, xi) * lin(iy, eta) * lin(iz, zeta)
            dN[g, a, 0] = dlin(ix) * lin(iy, eta) * lin(iz, zeta) * 2.0 / hx
            dN[g, a, 1] = lin(ix, xi) * dlin(iy) * lin(iz, zeta) * 2.0 / hy
            dN[g, a, 2] = lin(ix, xi) * lin(iy, eta) * dlin(iz) * 2.0 / hz
    return N, dN


def elements(dims):
    nx, ny, nz = dims
    for k in range(nz - 1):
        for j in range(ny - 1):
            for i in range(nx - 1):
                yield i, j, k


def node_ids(dims, i, j, k):
    nx, ny, nz = dims
    return [(i + ix) + nx * ((j + iy) + ny * (k + iz)) for (ix, iy, iz) in LOCAL]


def gp_values(field, dims, N):
    """Gauss-point interpolation with the zero-node rule: an element whose 8
    nodal values contain an exact 0.0 gets 0 at all its Gauss points."""
    flat = np.asarray(field).reshape(-1)
    out = {}
    for (i, j, k) in elements(dims):
        v = flat[node_ids(dims, i, j, k)]
        out[(i, j, k)] = np.zeros(8) if np.any(v == 0.0) else N @ v
    return out


def degrade(dg, k):
    dc = np.clip(dg, 0.0, 1.0)
    return (1.0 - k) * (1.0 - dc) ** 2 + k


def fine_coefficients(dims, h, Kn, Gn, dn, flag, small_k):
    """Gauss-point (K, G) of the finest level: K degraded when flag == 0, G
    degraded unless flag == 2 (the split-flag rule of the reference model)."""
    N, _ = hex8_tables(h)
    Kg, Gg, dg = gp_values(Kn, dims, N), gp_values(Gn, dims, N), gp_values(dn, dims, N)
    fl = np.asarray(flag).reshape(-1)
    nx, ny, _ = dims
    coef = {}
    for e in elements(dims):
        q = e[0] + nx * (e[1] + ny * e[2])
        gfac = degrade(dg[e], small_k)
        Kd = Kg[e] * gfac if fl[q] == 0 else Kg[e]
        Gd = Gg[e] if fl[q] == 2 else Gg[e] * gfac
        coef[e] = (Kd, Gd)
    return coef


def coarse_coefficients(dims, Ke, Ge):
    nx, ny, _ = dims
    Kf, Gf = np.asarray(Ke).reshape(-1), np.asarray(Ge).reshape(-1)
    return {e: (np.full(8, Kf[e[0] + nx * (e[1] + ny * e[2])]), np.full(8, Gf[e[0] + nx * (e[1] + ny * e[2])]))
            for e in elements(dims)}


def bmatrix(dN_g):
    """6x24 Voigt strain-displacement matrix, DOF order (a, c) -> 3a + c."""
    B = np.zeros((6, 24))
    for a in range(8):
        dx, dy, dz = dN_g[a]
        B[0, 3 * a + 0] = dx
        B[1, 3 * a + 1] = dy
        B[2, 3 * a + 2] = dz
        B[3, 3 * a + 0] = dy
        B[3, 3 * a + 1] = dx
        B[4, 3 * a + 1] = dz
        B[4, 3 * a + 2] = dy
        B[5, 3 * a + 0] = dz
        B[5, 3 * a + 2] = dx
    return B


def dmatrix(K, G):
    lam = K - 2.0 * G / 3.0
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[0, 0] = D[1, 1] = D[2, 2] = lam + 2 * G
    D[3, 3] = D[4, 4] = D[5, 5] = G
    return D


def elastic_matrix(dims, h, coef):
    """Assembled degraded-elasticity stiffness, global DOF index c*Nn + node."""
    _, dN = hex8_tables(h)
    Bs = [bmatrix(dN[g]) for g in range(8)]
    wdet = h[0] * h[1] * h[2] / 8.0
    Nn = int(np.prod(dims))
    rows, cols, vals = [], [], []
    for e in elements(dims):
        Kg, Gg = coef[e]
        Ke = np.zeros((24, 24))
        for g in range(8):
            Ke += wdet * Bs[g].T @ dmatrix(Kg[g], Gg[g]) @ Bs[g]
        nid = node_ids(dims, *e)
        gl = [c * Nn + nid[a] for a in range(8) for c in range(3)]
        r, cc = np.meshgrid(gl, gl, indexing="ij")
        rows.append(r.ravel())
        cols.append(cc.ravel())
        vals.append(Ke.ravel())
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(3 * Nn, 3 * Nn))


def phase_matrix(dims, h, H_elem, gc_nodal, lc):
    """Assembled AT2 operator: (2H + gc/lc) mass + gc*lc stiffness."""
    N, dN = hex8_tables(h)
    gcg = gp_values(gc_nodal, dims, N)
    Hf = np.asarray(H_elem).reshape(-1)
    nx, ny, _ = dims
    wdet = h[0] * h[1] * h[2] / 8.0
    Nn = int(np.prod(dims))
    rows, cols, vals = [], [], []
    for e in elements(dims):
        He = Hf[e[0] + nx * (e[1] + ny * e[2])]
        Ae = np.zeros((8, 8))
        for g in range(8):
            Ae += wdet * ((2 * He + gcg[e][g] / lc) * np.outer(N[g], N[g]) + gcg[e][g] * lc * dN[g] @ dN[g].T)
        nid = node_ids(dims, *e)
        r, cc = np.meshgrid(nid, nid, indexing="ij")
        rows.append(r.ravel())
        cols.append(cc.ravel())
        vals.append(Ae.ravel())
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(Nn, Nn))


def interp1d_matrix(nc):
    """Linear interpolation from nc coarse nodes to 2(nc-1)+1 fine nodes."""
    nf = 2 * (nc - 1) + 1
    P = np.zeros((nf, nc))
    for i in range(nf):
        if i % 2 == 0:
            P[i, i // 2] = 1.0
        else:
            P[i, i // 2] = 0.5
            P[i, i // 2 + 1] = 0.5
    return P


def prolongation(dims_c):
    """Trilinear prolongation as a Kronecker product (x fastest ordering)."""
    nx, ny, nz = dims_c
    return sp.csr_matrix(np.kron(interp1d_matrix(nz), np.kron(interp1d_matrix(ny), interp1d_matrix(nx))))


def injection(dims_c):
    nx, ny, nz = dims_c
    def inj(n):
        m = np.zeros((n, 2 * (n - 1) + 1))
        for i in range(n):
            m[i, 2 * i] = 1.0
        return m
    return sp.csr_matrix(np.kron(inj(nz), np.kron(inj(ny), inj(nx))))


def dirichlet(dims, h, origin, domain_end, Ut):
    """Dirichlet DOFs and values of the reference's boundary conditions:
    u_z = 0 on the bottom face, u_z = Ut on the top face, u_x = 0 at the
    bottom node (x_mid, y_a) and u_y = 0 at the bottom node (x_b, y_mid),
    all located by coordinate equality."""
    nx, ny, nz = dims
    xa, ya, za = origin
    xb, yb, zb = domain_end
    Nn = nx * ny * nz
    vals = {}
    for k in range(nz):
        z = za + k * h[2]
        for j in range(ny):
            y = ya + j * h[1]
            for i in range(nx):
                x = xa + i * h[0]
                q = i + nx * (j + ny * k)
                if z == za and x == 0.5 * (xa + xb) and y == ya:
                    vals[0 * Nn + q] = 0.0
                if z == za and x == xb and y == 0.5 * (ya + yb):
                    vals[1 * Nn + q] = 0.0
                if z == za:
                    vals[2 * Nn + q] = 0.0
                if z == zb:
                    vals[2 * Nn + q] = Ut
    return vals


def jacobi_pcg(A, b, x0, nb, free=None):
    """Textbook Jacobi-preconditioned CG, nb iterations from x0, on the rows
    and columns listed in `free` (all when None).  Returns the iterate and the
    final residual (zero on constrained rows)."""
    n = A.shape[0]
    free = np.arange(n) if free is None else np.asarray(free)
    Aff = A[free][:, free]
    x = x0.copy()
    r_full = b - A @ x
    r = r_full[free]
    dinv = 1.0 / Aff.diagonal()
    p = np.zeros_like(r)
    rz_old = None
    for it in range(nb):
        z = dinv * r
        rz = r @ z
        beta = 0.0 if it == 0 else rz / rz_old
        p = z + beta * p
        w = Aff @ p
        alpha = rz / (p @ w)
        x[free] += alpha * p
        r = r - alpha * w
        rz_old = rz
    rr = np.zeros(n)
    rr[free] = r
    return x, rr


def mesh(nodes_per_axis, seed, zero_frac=0.0, flags=True, dscale=0.6):
    """Random nodal K, G, d and element flags for operator tests."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = nodes_per_axis
    shape = (nz, ny, nx)
    K = rng.uniform(1.0, 3.0, shape)
    G = rng.uniform(0.5, 1.5, shape)
    d = rng.uniform(-0.1, 1.1, shape) * dscale
    if zero_frac > 0:
        for arr in (K, G, d):
            m = rng.random(shape) < zero_frac
            arr[m] = 0.0
    fl = rng.integers(0, 3, shape).astype(np.float64) if flags else np.zeros(shape)
    return K, G, d, fl


def all_combos(n):
    return list(itertools.product(range(n), repeat=3))
