"""Python mirror of HPCrack's solve entry points over the C-ABI.

`Solver` holds one device-resident problem (an hpcr_ctx).  Its methods keep
the reference's names and argument meaning -- u_entire, phase_field, pcg_u,
pcg_d, cycle_u, cycle_d, Internal_Force, PreM_u, Fld, Precondition_M,
strain_history, U_e, ... -- and raise HpcrError where the reference would
exit or throw.  Arrays are numpy views in (z, y, x) order, i.e. x-fastest
like the reference's grid<T> (include/grid.h:85).
"""
import ctypes

import numpy as np

from . import capi
from .capi import (D, DOLD, DPREV, FD, FLAG, FU, G, GC, GE, H, K, KE, P, PD, RD, RU, U, UOLD, W, WD,
                   VECTOR_FIELDS)


class HpcrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hpcr error {code}: {msg}")
        self.code = code


def device_count():
    n = ctypes.c_int(0)
    rc = capi.load().hpcr_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


def estimate_bytes(maxlevel, element_nb, domain, mixed=False):
    """Device bytes a Solver of this geometry holds (no device needed);
    mixed: True (or 1) adds the mixed-precision solve's buffers, 2 those of
    its FP64-preconditioner form."""
    L = capi.load()
    geo = capi.Geometry()
    geo.maxlevel = maxlevel
    for i in range(3):
        geo.element_nb[i] = int(element_nb[i])
    geo.xa, geo.xb, geo.ya, geo.yb, geo.za, geo.zb = [float(v) for v in domain]
    out = ctypes.c_longlong(0)
    rc = L.hpcr_estimate_bytes(ctypes.byref(geo), int(mixed), ctypes.byref(out))
    if rc != 0:
        raise HpcrError(rc, "bad geometry")
    return out.value


class Fuser:
    """Fusion engine (hpcr_fuser): the multigrid cycles of the Solvers
    attached to it run as multi-problem launches on one stream, each problem
    over its own grid (bit-identical results).  Solves of different Solvers
    may run from different threads (ctypes releases the GIL)."""

    def __init__(self, window_us=200, graphs=True, device=0):
        L = capi.load()
        self._L = L
        h = ctypes.c_void_p()
        rc = L.hpcr_fuser_create(device, int(window_us), ctypes.byref(h))
        if rc != 0:
            raise HpcrError(rc, "cannot create the fuser")
        self._h = h
        if not graphs:
            L.hpcr_fuser_set_graphs(h, 0)

    def stats(self):
        st = capi.FuserStats()
        rc = self._L.hpcr_fuser_get_stats(self._h, ctypes.byref(st))
        if rc != 0:
            raise HpcrError(rc, "fuser stats")
        return {k: getattr(st, k) for k, _ in capi.FuserStats._fields_}

    def close(self):
        """Destroys the fuser; every attached Solver must be detached or closed first."""
        if getattr(self, "_h", None):
            rc = self._L.hpcr_fuser_destroy(self._h)
            if rc != 0:
                raise HpcrError(rc, "fuser still has attached contexts")
            self._h = None


class Solver:
    def __init__(self, maxlevel, element_nb, domain, small_k=1e-5, lc=3.0,
                 mg_u=(50, 4, 2, 1, 12), mg_d=(10, 2, 1, 1, 12),
                 tol_u=1e-6, tol_d_rel=1e-12, tol_d_abs=1e-12, device=0):
        L = capi.load()
        self._L = L
        geo = capi.Geometry()
        geo.maxlevel = maxlevel
        for i in range(3):
            geo.element_nb[i] = int(element_nb[i])
        geo.xa, geo.xb, geo.ya, geo.yb, geo.za, geo.zb = [float(v) for v in domain]
        par = capi.Params()
        par.small_k, par.lc, par.tol_u, par.tol_d_rel, par.tol_d_abs = small_k, lc, tol_u, tol_d_rel, tol_d_abs
        par.mg_u = capi.MG(*mg_u)
        par.mg_d = capi.MG(*mg_d)
        h = ctypes.c_void_p()
        rc = L.hpcr_ctx_create(ctypes.byref(geo), ctypes.byref(par), device, ctypes.byref(h))
        if rc != 0:
            raise HpcrError(rc, L.hpcr_last_error(None).decode())
        self._h = h
        self.maxlevel = maxlevel
        self.small_k, self.lc = small_k, lc
        self.tol_u = tol_u
        self.last_stats = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.hpcr_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise HpcrError(rc, self._L.hpcr_last_error(self._h).decode())

    # ---- geometry / fields ------------------------------------------------
    def dims(self, l):
        d = (ctypes.c_int * 3)()
        self._chk(self._L.hpcr_level_dims(self._h, l, d))
        return tuple(d)

    def h(self, l):
        d = (ctypes.c_double * 3)()
        self._chk(self._L.hpcr_level_h(self._h, l, d))
        return tuple(d)

    def device_bytes(self):
        return self._L.hpcr_device_bytes(self._h)

    def get(self, f, l, comp=None):
        nx, ny, nz = self.dims(l)
        comps = [comp] if comp is not None else (range(3) if f in VECTOR_FIELDS else [0])
        out = []
        for c in comps:
            a = np.empty((nz, ny, nx), dtype=np.float64)
            self._chk(self._L.hpcr_get_field(self._h, f, l, c, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
            out.append(a)
        return np.stack(out) if (comp is None and f in VECTOR_FIELDS) else out[0]

    def set(self, f, l, arr, comp=None):
        nx, ny, nz = self.dims(l)
        arr = np.asarray(arr, dtype=np.float64)
        if comp is None and f in VECTOR_FIELDS:
            arr = np.broadcast_to(arr, (3, nz, ny, nx))
            for c in range(3):
                a = np.ascontiguousarray(arr[c])
                self._chk(self._L.hpcr_set_field(self._h, f, l, c, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        else:
            a = np.ascontiguousarray(np.broadcast_to(arr, (nz, ny, nx)))
            self._chk(self._L.hpcr_set_field(self._h, f, l, comp or 0,
                                             a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def fill(self, f, l, value, comp=0):
        self._chk(self._L.hpcr_fill_field(self._h, f, l, comp, float(value)))

    def synchronize(self):
        self._chk(self._L.hpcr_synchronize(self._h))

    # ---- set-up --------------------------------------------------------------
    def coarse_vr(self, f):
        self._chk(self._L.hpcr_coarse_vr(self._h, f))

    # ---- operators -------------------------------------------------------------
    def internal_force(self, l, src=P, dst=W):
        self._chk(self._L.hpcr_internal_force(self._h, l, src, dst))

    def internal_force_f32(self, src=P, dst=W, mfma=False):
        """The mixed solve's FP32 finest-level operator: dst = -(A src),
        zero on Dirichlet DOFs (matrix-core element contraction with mfma)."""
        self._chk(self._L.hpcr_internal_force_f32(self._h, src, dst, 1 if mfma else 0))

    def prem_u(self, l, dst=W):
        self._chk(self._L.hpcr_prem_u(self._h, l, dst))

    def fld(self, l, src=PD, dst=WD):
        self._chk(self._L.hpcr_fld(self._h, l, src, dst))

    def precondition_m(self, l, dst=WD):
        self._chk(self._L.hpcr_precondition_m(self._h, l, dst))

    def fd_h(self, l):
        self._chk(self._L.hpcr_fd_h(self._h, l))

    def mp_choice(self):
        self._chk(self._L.hpcr_mp_choice(self._h))

    def mp_all_element(self):
        self._chk(self._L.hpcr_mp_all_element(self._h))

    def mech_setup(self):
        self._chk(self._L.hpcr_mech_setup(self._h))

    def vr_all_level(self, f):
        self._chk(self._L.hpcr_vr_all_level(self._h, f))

    def voigt_all_level(self, f):
        self._chk(self._L.hpcr_voigt_all_level(self._h, f))

    def dirichlet_bd(self, l, f, Ut):
        self._chk(self._L.hpcr_dirichlet_bd(self._h, l, f, Ut))

    def injection(self, lc, f, comp=0):
        self._chk(self._L.hpcr_injection(self._h, lc, f, comp))

    def restrict(self, lf, rf, fic, dst, comp=0):
        self._chk(self._L.hpcr_restrict(self._h, lf, rf, fic, dst, comp))

    def prolong(self, lf, uf, uc, ucold, comp=0):
        self._chk(self._L.hpcr_prolong(self._h, lf, uf, uc, ucold, comp))

    # ---- smoothers, cycles, solves ------------------------------------------------
    def pcg_u(self, l, nb, Ut, t, f_norm=0.0):
        fn = ctypes.c_double(f_norm)
        self._chk(self._L.hpcr_pcg_u(self._h, l, nb, Ut, t, ctypes.byref(fn)))
        return fn.value

    def pcg_d(self, l, nb, fd_norm=0.0):
        fn = ctypes.c_double(fd_norm)
        self._chk(self._L.hpcr_pcg_d(self._h, l, nb, ctypes.byref(fn)))
        return fn.value

    def cycle_u(self, l, Ut, t, f_norm=0.0):
        fn = ctypes.c_double(f_norm)
        self._chk(self._L.hpcr_cycle_u(self._h, l, Ut, t, ctypes.byref(fn)))
        return fn.value

    def cycle_d(self, l, fd_norm=0.0):
        fn = ctypes.c_double(fd_norm)
        self._chk(self._L.hpcr_cycle_d(self._h, l, ctypes.byref(fn)))
        return fn.value

    def u_entire(self, Ut, t, f_norm):
        """Mechanical solve; returns (f_norm, cycles)."""
        fn = ctypes.c_double(f_norm)
        st = capi.Stats()
        self._chk(self._L.hpcr_u_entire(self._h, Ut, t, ctypes.byref(fn), ctypes.byref(st)))
        self.last_stats = st
        return fn.value, st.cycles

    def set_fuser(self, fuser):
        """Run this Solver's multigrid cycles through `fuser` (None detaches)."""
        self._chk(self._L.hpcr_set_fuser(self._h, fuser._h if fuser is not None else None))

    def set_reduction_fence(self, on):
        """Order reduction partials by a release/acquire fence pair instead
        of a completion wait (identical results)."""
        self._chk(self._L.hpcr_set_reduction_fence(self._h, 1 if on else 0))

    def set_graphs(self, on):
        """Replay whole multigrid cycles as HIP graphs (identical results)."""
        self._chk(self._L.hpcr_set_graphs(self._h, 1 if on else 0))

    def u_entire_mixed(self, Ut, t, f_norm, tol=None, max_iter=100, cheb_degree=0, power_iters=10,
                       cheb_ratio=6.0, cheb_upper=1.1, replace_every=10, precond_fp64=False,
                       element_mfma=False, coarse_iters=0):
        """Mixed-precision form of u_entire (FP64 flexible CG, FP32 V-cycle);
        returns (f_norm, outer iterations).  tol defaults to the context's
        tol_u (squared relative residual).  cheb_degree 0 chooses the
        Chebyshev degree, range and safety factor from the modulus contrast
        (mixed_resolve); otherwise cheb_ratio and cheb_upper apply.
        precond_fp64: the same V-cycle in FP64 (the like-for-like comparison
        of the two precisions).  element_mfma: the FP32 finest-level element
        contraction on the matrix cores.  coarse_iters: iterations of the
        preconditioner's coarsest solve (0: the context's nu0_u)."""
        fn = ctypes.c_double(f_norm)
        st = capi.Stats()
        mp = capi.MixedParams(max_iter, self.tol_u if tol is None else tol, cheb_degree, power_iters, cheb_ratio,
                              cheb_upper, replace_every, 1 if precond_fp64 else 0, 1 if element_mfma else 0,
                              coarse_iters)
        self._chk(self._L.hpcr_u_entire_mixed(self._h, Ut, t, ctypes.byref(fn), ctypes.byref(mp), ctypes.byref(st)))
        self.last_stats = st
        return fn.value, st.cycles

    def mixed_resolve(self, cheb_degree=0, cheb_ratio=6.0, cheb_upper=1.1):
        """(cheb_degree, cheb_ratio, cheb_upper, contrast) that u_entire_mixed
        uses on the current fields; contrast is 0 unless cheb_degree is 0."""
        mp = capi.MixedParams(100, self.tol_u, cheb_degree, 10, cheb_ratio, cheb_upper, 10, 0, 0, 0)
        k = ctypes.c_double(0.0)
        self._chk(self._L.hpcr_mixed_resolve(self._h, ctypes.byref(mp), ctypes.byref(k)))
        return mp.cheb_degree, mp.cheb_ratio, mp.cheb_upper, k.value

    def phase_field(self, t, fd_norm):
        """Phase-field solve; returns (fd_norm, cycles)."""
        fn = ctypes.c_double(fd_norm)
        st = capi.Stats()
        self._chk(self._L.hpcr_phase_field(self._h, t, ctypes.byref(fn), ctypes.byref(st)))
        self.last_stats = st
        return fn.value, st.cycles

    def history(self):
        st = self.last_stats
        return np.array([st.r_norm_initial] + list(st.history[:min(st.cycles, capi.HIST_MAX)]))

    def strain_history(self, u_prop):
        self._chk(self._L.hpcr_strain_history(self._h, u_prop))

    def _scalar(self, fn):
        v = ctypes.c_double(0.0)
        self._chk(fn(self._h, ctypes.byref(v)))
        return v.value

    def U_e(self):
        return self._scalar(self._L.hpcr_U_e)

    def externe_force_sum(self):
        return self._scalar(self._L.hpcr_externe_force_sum)

    def delta_d_max(self):
        return self._scalar(self._L.hpcr_delta_d_max)

    def d_max(self):
        return self._scalar(self._L.hpcr_d_max)

    def save_d_old(self):
        self._chk(self._L.hpcr_save_d_old(self._h))

    # ---- measurement ----------------------------------------------------------------
    def timing(self, on=True):
        self._chk(self._L.hpcr_timing_enable(self._h, 1 if on else 0))

    def timing_reset(self):
        self._chk(self._L.hpcr_timing_reset(self._h))

    def timing_query(self, kernel, level):
        ms = ctypes.c_double(0.0)
        n = ctypes.c_longlong(0)
        self._chk(self._L.hpcr_timing_query(self._h, kernel, level, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
