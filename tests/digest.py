"""Compact digests of large fields, for golden fixtures of the oracle at
sizes whose full fields are too big to commit (129^3 and 257^3 nodes).

A digest holds the values at a fixed, seeded sample of nodes, the sums over
every z-, y- and x-plane, and the squared norm.  A field is compared with a
digest through the sample (relative L2 over the sampled values) and through
the plane sums; for displacements both first remove the least-squares
amount of the zero-energy rotation mode from the difference (the
reference's boundary conditions leave that mode free, so its amount is set
by rounding)."""
import numpy as np

from staggered import rotation_mode

N_SAMPLE = 4096


def sample_nodes(dims, n=N_SAMPLE, seed=12345):
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(nx * ny * nz, size=min(n, nx * ny * nz), replace=False))


def _comps(a, dims):
    nx, ny, nz = dims
    a = np.asarray(a, dtype=np.float64)
    return a.reshape(-1, nz, ny, nx)


def _plane_sums(c):
    # c: (ncomp, nz, ny, nx) -> concatenated sums over z-, y- and x-planes
    return np.concatenate([np.concatenate([x.sum(axis=(1, 2)), x.sum(axis=(0, 2)), x.sum(axis=(0, 1))]) for x in c])


def make(a, dims):
    c = _comps(a, dims)
    idx = sample_nodes(dims)
    flat = c.reshape(c.shape[0], -1)
    return dict(sample=flat[:, idx].copy(), planes=_plane_sums(c), norm2=np.array([float(np.sum(c * c))]))


def compare(a, dg, dims, domain=None):
    """(relative L2 over the sample, relative max over the plane sums) of
    field a against digest dg; with domain given, a is a displacement and
    the rotation mode is removed from the difference first."""
    c = _comps(a, dims)
    idx = sample_nodes(dims)
    flat = c.reshape(c.shape[0], -1)
    es = (flat[:, idx] - dg["sample"]).reshape(-1)
    ep = _plane_sums(c) - dg["planes"]
    ref_s = np.asarray(dg["sample"]).reshape(-1)
    ref_p = np.asarray(dg["planes"])
    if domain is not None:
        m = rotation_mode(dims, domain).reshape(c.shape[0], -1)
        ms = m[:, idx].reshape(-1)
        es = es - (es @ ms) / (ms @ ms) * ms
        mp = _plane_sums(m.reshape(c.shape))
        ep = ep - (ep @ mp) / (mp @ mp) * mp
    s = np.linalg.norm(es) / max(np.linalg.norm(ref_s), 1e-300)
    p = np.max(np.abs(ep)) / max(np.max(np.abs(ref_p)), 1e-300)
    return float(s), float(p)
