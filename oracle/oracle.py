"""ctypes binding of the oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
module; the product (hpcrack_amd) never does.  The oracle is the CPU
restatement of HPCrack's MG/PCG hot path in hpcr_oracle.cpp.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libhpcr_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_pointwise.so")

# Field ids shared with hpcrack_amd (include/hpcrack.h, enum hpcr_field)
U, FU, RU, UOLD, P, W, D, FD, RD, DOLD, H, K, G, GC, KE, GE, FLAG, DPREV, PD, WD = range(20)
NCOMP = {U: 3, FU: 3, RU: 3, UOLD: 3, P: 3, W: 3}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        vp = ctypes.c_void_p
        L.ora_create.restype = vp
        L.ora_create.argtypes = [ctypes.c_int, ip, dp, ctypes.c_double, ctypes.c_double, ip, ip,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.ora_destroy.argtypes = [vp]
        L.ora_error.restype = ctypes.c_char_p
        L.ora_error.argtypes = [vp]
        L.ora_dims.argtypes = [vp, ctypes.c_int, ip]
        L.ora_h.argtypes = [vp, ctypes.c_int, dp]
        L.ora_field.restype = dp
        L.ora_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_long)]
        L.ora_history.restype = ctypes.c_int
        L.ora_history.argtypes = [vp, ctypes.c_int, dp, ctypes.c_int]
        L.ora_shape_functions.argtypes = [ctypes.c_double] * 3 + [dp, dp]
        for n in ("ora_internal_force", "ora_fld"):
            getattr(L, n).argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        for n in ("ora_prem_u", "ora_precondition_m", "ora_dirichlet_residual"):
            getattr(L, n).argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.ora_fd_h.argtypes = [vp, ctypes.c_int]
        for n in ("ora_mp_choice", "ora_mp_all_element", "ora_mech_setup", "ora_save_d_old"):
            getattr(L, n).argtypes = [vp]
        for n in ("ora_vr_all_level", "ora_voigt_all_level", "ora_coarse_vr_nodal"):
            getattr(L, n).argtypes = [vp, ctypes.c_int]
        L.ora_dirichlet_bd.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.ora_injection.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_restrict.argtypes = [vp] + [ctypes.c_int] * 5
        L.ora_prolong.argtypes = [vp] + [ctypes.c_int] * 5
        L.ora_pcg_u.restype = ctypes.c_int
        L.ora_pcg_u.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp]
        L.ora_pcg_d.restype = ctypes.c_int
        L.ora_pcg_d.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.ora_cycle_u.restype = ctypes.c_int
        L.ora_cycle_u.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp]
        L.ora_cycle_d.restype = ctypes.c_int
        L.ora_cycle_d.argtypes = [vp, ctypes.c_int, dp]
        L.ora_u_entire.restype = ctypes.c_int
        L.ora_u_entire.argtypes = [vp, ctypes.c_double, ctypes.c_int, dp, ip]
        L.ora_phase_field.restype = ctypes.c_int
        L.ora_phase_field.argtypes = [vp, ctypes.c_int, dp, ip]
        L.ora_strain_history.argtypes = [vp, ctypes.c_double]
        for n in ("ora_U_e", "ora_externe_force_sum", "ora_delta_d_max", "ora_d_max"):
            getattr(L, n).restype = ctypes.c_double
            getattr(L, n).argtypes = [vp]
        L.ora_image_mp.restype = ctypes.c_int
        L.ora_image_mp.argtypes = [vp, ctypes.POINTER(ctypes.c_ubyte), ip, ctypes.c_int, dp, dp, dp,
                                   ctypes.c_double]
        L.ora_set_threads.argtypes = [ctypes.c_int]
        L.ora_set_ref_loops.argtypes = [ctypes.c_int]
        L.ora_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def shape_functions(hx, hy, hz):
    N = np.zeros(64)
    DN = np.zeros(192)
    lib().ora_shape_functions(hx, hy, hz, _dp(N), _dp(DN))
    return N.reshape(8, 8), DN.reshape(3, 8, 8)


def set_threads(n):
    lib().ora_set_threads(int(n))


def set_ref_loops(on):
    """Internal_Force and PreM_u in the reference's loop form (redundant
    Gauss-point coefficients, critical-section scatter): for timing the CPU
    baseline only; the scatter order, and so the rounding, then depends on
    thread timing."""
    lib().ora_set_ref_loops(1 if on else 0)


class Oracle:
    """One problem hierarchy on the CPU (mirrors a Stack plus its mg fields)."""

    def __init__(self, maxlevel, element_nb, domain, small_k=1e-5, lc=3.0,
                 mg_u=(50, 4, 2, 1, 12), mg_d=(10, 2, 1, 1, 12),
                 tol_u=1e-6, tol_d_rel=1e-12, tol_d_abs=1e-12):
        L = lib()
        self.maxlevel = maxlevel
        en = np.array(element_nb, dtype=np.int32)
        dom = np.array(domain, dtype=np.float64)
        mu = np.array(mg_u, dtype=np.int32)
        md = np.array(mg_d, dtype=np.int32)
        self._h = L.ora_create(maxlevel, _ip(en), _dp(dom), small_k, lc, _ip(mu), _ip(md),
                               tol_u, tol_d_rel, tol_d_abs)
        self.small_k, self.lc = small_k, lc

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ora_destroy(self._h)
            self._h = None

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(lib().ora_error(self._h).decode())

    def dims(self, l):
        out = np.zeros(3, dtype=np.int32)
        lib().ora_dims(self._h, l, _ip(out))
        return tuple(int(v) for v in out)

    def h(self, l):
        out = np.zeros(3)
        lib().ora_h(self._h, l, _dp(out))
        return tuple(out)

    def field(self, f, l, comp=0):
        """Writable numpy view (z, y, x) of a field on level l."""
        nx, ny, nz = self.dims(l)
        ptr = lib().ora_field(self._h, f, l, comp)
        return np.ctypeslib.as_array(ptr, shape=(nz, ny, nx))

    def vec(self, f, l):
        return np.stack([self.field(f, l, c) for c in range(NCOMP.get(f, 1))])

    def set(self, f, l, arr, comp=None):
        arr = np.asarray(arr, dtype=np.float64)
        if comp is None and NCOMP.get(f, 1) == 3:
            for c in range(3):
                self.field(f, l, c)[...] = arr[c]
        else:
            self.field(f, l, comp or 0)[...] = arr

    def stats(self):
        out = np.zeros(4, dtype=np.int64)
        lib().ora_stats(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        return dict(pcg_u_iters=int(out[0]), pcg_d_iters=int(out[1]),
                    applies_u=int(out[2]), applies_d=int(out[3]))

    def history(self, which):
        buf = np.zeros(256)
        n = lib().ora_history(self._h, 0 if which == "u" else 1, _dp(buf), 256)
        return buf[:n].copy()

    # operator-level
    def internal_force(self, l, src=P, dst=W):
        lib().ora_internal_force(self._h, l, src, dst)

    def prem_u(self, l, dst=W):
        lib().ora_prem_u(self._h, l, dst)

    def fld(self, l, src=PD, dst=WD):
        lib().ora_fld(self._h, l, src, dst)

    def precondition_m(self, l, dst=WD):
        lib().ora_precondition_m(self._h, l, dst)

    def fd_h(self, l):
        lib().ora_fd_h(self._h, l)

    def mp_choice(self):
        lib().ora_mp_choice(self._h)

    def mp_all_element(self):
        lib().ora_mp_all_element(self._h)

    def vr_all_level(self, f):
        lib().ora_vr_all_level(self._h, f)

    def voigt_all_level(self, f):
        lib().ora_voigt_all_level(self._h, f)

    def coarse_vr_nodal(self, f):
        lib().ora_coarse_vr_nodal(self._h, f)

    def mech_setup(self):
        lib().ora_mech_setup(self._h)

    def dirichlet_bd(self, l, f, Ut):
        lib().ora_dirichlet_bd(self._h, l, f, Ut)

    def dirichlet_residual(self, l, f):
        lib().ora_dirichlet_residual(self._h, l, f)

    def injection(self, lc, f, comp=0):
        lib().ora_injection(self._h, lc, f, comp)

    def restrict(self, lf, rf, fic, dst, comp=0):
        lib().ora_restrict(self._h, lf, rf, fic, dst, comp)

    def prolong(self, lf, uf, uc, ucold, comp=0):
        lib().ora_prolong(self._h, lf, uf, uc, ucold, comp)

    # smoothers, cycles, solves
    def pcg_u(self, l, nb, Ut, t, f_norm=0.0):
        fn = ctypes.c_double(f_norm)
        self._chk(lib().ora_pcg_u(self._h, l, nb, Ut, t, ctypes.byref(fn)))
        return fn.value

    def pcg_d(self, l, nb, fd_norm=0.0):
        fn = ctypes.c_double(fd_norm)
        self._chk(lib().ora_pcg_d(self._h, l, nb, ctypes.byref(fn)))
        return fn.value

    def cycle_u(self, l, Ut, t, f_norm=0.0):
        fn = ctypes.c_double(f_norm)
        self._chk(lib().ora_cycle_u(self._h, l, Ut, t, ctypes.byref(fn)))
        return fn.value

    def cycle_d(self, l, fd_norm=0.0):
        fn = ctypes.c_double(fd_norm)
        self._chk(lib().ora_cycle_d(self._h, l, ctypes.byref(fn)))
        return fn.value

    def u_entire(self, Ut, t, f_norm):
        fn = ctypes.c_double(f_norm)
        cy = ctypes.c_int(0)
        self._chk(lib().ora_u_entire(self._h, Ut, t, ctypes.byref(fn), ctypes.byref(cy)))
        return fn.value, cy.value

    def phase_field(self, t, fd_norm):
        fn = ctypes.c_double(fd_norm)
        cy = ctypes.c_int(0)
        self._chk(lib().ora_phase_field(self._h, t, ctypes.byref(fn), ctypes.byref(cy)))
        return fn.value, cy.value

    def strain_history(self, u_prop):
        lib().ora_strain_history(self._h, u_prop)

    def U_e(self):
        return lib().ora_U_e(self._h)

    def externe_force_sum(self):
        return lib().ora_externe_force_sum(self._h)

    def delta_d_max(self):
        return lib().ora_delta_d_max(self._h)

    def d_max(self):
        return lib().ora_d_max(self._h)

    def save_d_old(self):
        lib().ora_save_d_old(self._h)

    def image_mp(self, img, gray, E, nu, gc, voxel_size):
        img = np.ascontiguousarray(img, dtype=np.uint8)
        gray = np.ascontiguousarray(gray, dtype=np.int32)
        E = np.ascontiguousarray(E, dtype=np.float64)
        nu = np.ascontiguousarray(nu, dtype=np.float64)
        gc = np.ascontiguousarray(gc, dtype=np.float64)
        self._chk(lib().ora_image_mp(self._h, img.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                                     _ip(gray), len(gray), _dp(E), _dp(nu), _dp(gc), voxel_size))


def ref_lib():
    """The reference's MPI-free pointwise routines, or None when not built."""
    if not os.path.exists(REF_LIB):
        return None
    L = ctypes.CDLL(REF_LIB)
    dp = ctypes.POINTER(ctypes.c_double)
    L.ref_dshape_function.argtypes = [ctypes.c_double] * 3 + [dp, dp]
    L.ref_element_strain_stress.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, dp, dp, dp, dp]
    return L
