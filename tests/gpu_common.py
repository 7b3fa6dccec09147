"""Shared set-up of the GPU parity tests: one oracle problem and one HIP
problem built from the same seeded inputs."""
import sys

import numpy as np
import pytest

import oracle as O


def require_gpu():
    import hpcrack_amd as HA
    n = HA.device_count()
    if n < 1:
        pytest.skip("no HIP device visible")
    return HA


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


def spheres_labels(dims, n_incl, rmin, rmax, seed):
    """Random non-overlapping spherical inclusions (label 1) in matrix 0."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    lab = np.zeros((nz, ny, nx), dtype=np.uint8)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    placed = []
    tries = 0
    while len(placed) < n_incl and tries < 10000:
        tries += 1
        r = rng.uniform(rmin, rmax)
        c = np.array([rng.integers(0, nx), rng.integers(0, ny), rng.integers(0, nz)], dtype=float)
        if any(np.linalg.norm(c - p[0]) < r + p[1] + 1 for p in placed):
            continue
        placed.append((c, r))
        lab[(xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 <= r * r] = 1
    return lab


def material_from_labels(lab, voxel_size=1e-6):
    """Label -> K, G, gc as image_MP does for integer labels (two phases)."""
    E = np.where(lab == 1, 70e9, 3e9) * voxel_size
    nu = np.where(lab == 1, 0.2, 0.3)
    gc = np.where(lab == 1, 1000.0, 100.0)
    K = 1.0 / (3.0 * (1.0 - 2.0 * nu)) * E
    G = 1.0 / (2.0 * (1.0 + nu)) * E
    return K, G, gc


def make_pair(HA, maxlevel, enb, domain, seed=0, lab_incl=3, **kw):
    """Oracle + HIP solver with identical finest-level materials and coarse VR."""
    o = O.Oracle(maxlevel, enb, domain, **kw)
    s = HA.Solver(maxlevel, enb, domain, **kw)
    L = maxlevel
    dims = o.dims(L)
    assert dims == s.dims(L)
    lab = spheres_labels(dims, lab_incl, 1.5, max(2.0, min(dims) / 5.0), seed)
    K, G, gc = material_from_labels(lab)
    for f, a in ((O.K, K), (O.G, G), (O.GC, gc)):
        o.set(f, L, a)
        s.set(f, L, a)
    for f in (O.K, O.G, O.GC):
        o.coarse_vr_nodal(f)
        s.coarse_vr(f)
    return o, s, lab
