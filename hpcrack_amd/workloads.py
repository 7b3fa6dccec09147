"""Seeded synthetic workloads (voxel images = nodes, as in HPCrack where an
N^3-voxel image gives N^3 nodes).

Phases (gray level = label, so the stepwise CSV interpolation of image_MP is
exact): matrix 0 (E 3 GPa, nu 0.3, gc 0.1 N/mm = 100 N/m), stiff inclusion 1
(E 70 GPa, nu 0.2, gc 1 N/mm = 1000 N/m), pore 2 (E 1e-6 E_matrix, nu 0.3,
gc 1e-3 N/mm = 1 N/m).  Voxel 1 um, element_nb 4 on the coarsest level,
coordinates in voxel units (h_fine = 1), lc = 2 voxels, k = 1e-6.
"""
from dataclasses import dataclass, field

import numpy as np

from . import io as hio

GRAY = (0, 1, 2)
E_PA = (3e9, 70e9, 3e9 * 1e-6)
NU = (0.3, 0.2, 0.3)
GC = (100.0, 1000.0, 1.0)


@dataclass
class Workload:
    name: str
    n: int                       # nodes per axis
    seed: int
    n_incl: int
    incl_r: tuple
    pore_fraction: float = 0.0
    pore_r: tuple = (2.0, 5.0)
    element_nb: tuple = (4, 4, 4)
    voxel_size: float = 1e-6
    lc_voxels: float = 2.0
    small_k: float = 1e-6
    # squared relative residual: the reference tests squared norms
    # (pcg_utilities.cpp:35-58), so 1e-8 unsquared is 1e-16
    tol_u: float = 1e-16
    tol_d_rel: float = 1e-16
    tol_d_abs: float = 1e-30
    mg_u: tuple = (50, 4, 2, 1, 30)
    mg_d: tuple = (10, 2, 1, 1, 30)
    steps: int = 5
    # preset phase field: planar penny crack at mid-height, d = exp(-dist/l)
    penny_radius_frac: float = 0.0
    penny_l: float = 2.0
    # two-phase variant: inclusion Young's modulus (Pa) instead of the table's
    incl_E: float = 0.0
    # rejection-sampling budget of the sphere placement
    max_tries: int = 2000000
    extra: dict = field(default_factory=dict)

    @property
    def maxlevel(self):
        m, c = 0, self.element_nb[0]
        while 2 * c <= self.n - 1:
            c *= 2
            m += 1
        return m

    @property
    def domain(self):
        ext = [e * 2 ** self.maxlevel for e in self.element_nb]
        return (0.0, float(ext[0]), 0.0, float(ext[1]), 0.0, float(ext[2]))

    @property
    def dims(self):
        return tuple(e * 2 ** self.maxlevel + 1 for e in self.element_nb)

    def image(self):
        return hio.synth_image(self.dims, self.seed, n_incl=self.n_incl, incl_r=self.incl_r, incl_label=1,
                               pore_fraction=self.pore_fraction, pore_r=self.pore_r, pore_label=2,
                               max_tries=self.max_tries)

    def table(self):
        E = list(E_PA)
        if self.incl_E > 0:
            E[1] = self.incl_E
        return hio.material_table(GRAY, E, NU, GC)

    def damage(self):
        """Preset d (nz, ny, nx) of a planar penny crack {z = z_mid, rho <=
        R}, R = penny_radius_frac * (n - 1) voxels: d = exp(-dist / l) with
        dist the distance to the disk; None without a crack."""
        if self.penny_radius_frac <= 0:
            return None
        nx, ny, nz = self.dims
        R = self.penny_radius_frac * (self.n - 1)
        z = np.arange(nz, dtype=np.float64) - 0.5 * (nz - 1)
        y = np.arange(ny, dtype=np.float64) - 0.5 * (ny - 1)
        x = np.arange(nx, dtype=np.float64) - 0.5 * (nx - 1)
        rho = np.sqrt(y[:, None] ** 2 + x[None, :] ** 2)
        out = np.maximum(rho - R, 0.0)
        dist = np.sqrt(out[None, :, :] ** 2 + z[:, None, None] ** 2)
        return np.exp(-dist / self.penny_l)

    def materials(self):
        img, frac = self.image()
        K, G, gc = hio.image_mp(self.table(), img, self.voxel_size)
        return img, frac, K, G, gc

    def solver_kwargs(self):
        return dict(small_k=self.small_k, lc=self.lc_voxels, mg_u=self.mg_u, mg_d=self.mg_d, tol_u=self.tol_u,
                    tol_d_rel=self.tol_d_rel, tol_d_abs=self.tol_d_abs)

    def load_increment(self):
        """top u_z += 1e-4 * height per step"""
        return 1e-4 * (self.dims[2] - 1)


C1 = Workload("C1-129-two-phase", 129, 1001, n_incl=30, incl_r=(4.0, 12.0), steps=5)
C2 = Workload("C2-257-three-phase", 257, 1002, n_incl=200, incl_r=(3.0, 10.0), pore_fraction=0.05,
              pore_r=(2.0, 5.0), steps=20)
# 513^3 two-phase, inclusion/matrix stiffness ratio 50 (3 GPa / 150 GPa),
# inclusions to ~20 % volume (count chosen for that fraction at radii 4-12)
C3 = Workload("C3-513-two-phase-ratio50", 513, 1003, n_incl=17000, incl_r=(4.0, 12.0), incl_E=150e9, steps=3,
              max_tries=50000000)
# the C2 image with a preset penny crack (radius 0.3 L at mid-height, l = 2)
C4 = Workload("C4-257-penny-crack", 257, 1002, n_incl=200, incl_r=(3.0, 10.0), pore_fraction=0.05,
              pore_r=(2.0, 5.0), steps=1, penny_radius_frac=0.3, penny_l=2.0)
WORKLOADS = {"C1": C1, "C2": C2, "C3": C3, "C4": C4}


def scaled(w, n, name=None):
    """Same generator at another size, inclusion count scaled by volume."""
    f = ((n - 1) / (w.n - 1)) ** 3
    return Workload(name or f"{w.name}@{n}", n, w.seed, n_incl=max(1, int(round(w.n_incl * f))),
                    incl_r=w.incl_r, pore_fraction=w.pore_fraction, pore_r=w.pore_r, element_nb=w.element_nb,
                    voxel_size=w.voxel_size, lc_voxels=w.lc_voxels, small_k=w.small_k, tol_u=w.tol_u,
                    tol_d_rel=w.tol_d_rel, tol_d_abs=w.tol_d_abs, mg_u=w.mg_u, mg_d=w.mg_d, steps=w.steps,
                    penny_radius_frac=w.penny_radius_frac, penny_l=w.penny_l, incl_E=w.incl_E,
                    max_tries=w.max_tries)


def upload(solver, w, K, G, gc):
    """Finest-level materials, then coarse_VR on the device; the preset
    phase field if the workload has one."""
    L = w.maxlevel
    from .capi import D as F_D, GC as F_GC, G as F_G, K as F_K
    solver.set(F_K, L, K)
    solver.set(F_G, L, G)
    solver.set(F_GC, L, gc)
    for f in (F_K, F_G, F_GC):
        solver.coarse_vr(f)
    d = w.damage()
    if d is not None:
        solver.set(F_D, L, d)


def rotation_mode(w, level=None):
    """Flattened (3 x nodes) zero-energy mode of the reference's boundary
    conditions on a level: the rotation about the vertical axis through
    (x_b, y_a), which satisfies both pins and the u_z conditions.  Two
    solvers may leave different amounts of it in u."""
    L = w.maxlevel if level is None else level
    dims = tuple(e * 2 ** L + 1 for e in w.element_nb)
    dom = w.domain
    hx = (dom[1] - dom[0]) / (dims[0] - 1)
    hy = (dom[3] - dom[2]) / (dims[1] - 1)
    x = dom[0] + np.arange(dims[0]) * hx
    y = dom[2] + np.arange(dims[1]) * hy
    shape = (dims[2], dims[1], dims[0])
    mx = np.broadcast_to(-(y - dom[2])[None, :, None], shape)
    my = np.broadcast_to((x - dom[1])[None, None, :], shape)
    return np.concatenate([mx.ravel(), my.ravel(), np.zeros(int(np.prod(shape)))])


def label_counts(img):
    return {int(v): int(c) for v, c in zip(*np.unique(img, return_counts=True))}


def write_case(w, outdir):
    """The workload as input files of the C++ driver (hpcrack_run): the image
    as x-fastest uint8 img.raw, the material table as mat.csv (the reference's
    Material_Property.csv layout) and config.txt (the reference's config
    keys, config_reader order)."""
    import os
    os.makedirs(outdir, exist_ok=True)
    img, _ = w.image()
    hio.write_raw(os.path.join(outdir, "img.raw"), img)
    E = list(E_PA)
    if w.incl_E > 0:
        E[1] = w.incl_E
    with open(os.path.join(outdir, "mat.csv"), "w") as f:
        f.write("Gray level," + ",".join(str(g) for g in GRAY) + "\n")
        f.write("E/Pa," + ",".join(repr(float(e)) for e in E) + "\n")
        f.write("nu," + ",".join(repr(float(v)) for v in NU) + "\n")
        f.write("gc/(N/m)," + ",".join(repr(float(g)) for g in GC) + "\n")
    nx, ny, nz = w.dims
    mu, md = w.mg_u, w.mg_d
    with open(os.path.join(outdir, "config.txt"), "w") as f:
        f.write("D=3\nnp=1,1,1\nX_start=0,0,0\nIMGname=%s\n" % os.path.abspath(os.path.join(outdir, "img.raw")))
        f.write("voxel_nb_X=%d\nvoxel_nb_Y=%d\nvoxel_nb_Z=%d\n" % (nx, ny, nz))
        f.write("ROI_start_X=0\nROI_start_Y=0\nROI_start_Z=0\nvoxel_size=%r\n" % w.voxel_size)
        f.write("element_nb=%d,%d,%d\n" % tuple(w.element_nb))
        f.write("ncy_u=%d\nnu0_u=%d\nnu1_u=%d\nnu2_u=%d\ngamma_u=%d\n" % (mu[4], mu[0], mu[1], mu[2], mu[3]))
        f.write("ncy_d=%d\nnu0_d=%d\nnu1_d=%d\nnu2_d=%d\ngamma_d=%d\n" % (md[4], md[0], md[1], md[2], md[3]))
        f.write("lc=%d\n" % int(w.lc_voxels))
    return img


if __name__ == "__main__":
    import sys
    if len(sys.argv) != 4 or sys.argv[1] != "write-case" or sys.argv[2] not in WORKLOADS:
        print("usage: python -m hpcrack_amd.workloads write-case {%s} OUTDIR" % ",".join(WORKLOADS), file=sys.stderr)
        sys.exit(2)
    write_case(WORKLOADS[sys.argv[2]], sys.argv[3])
