"""The staggered load-step loop of crack_propagation (crack_propagation.cpp:
65-121) over an Oracle or a Solver (same method names), shared by the CPU
pins of the loop and the GPU parity tests."""
import numpy as np

import oracle as O


def staggered(obj, steps, voxel_size, lc, schedule="reference", du=None, log=None, unit=1.0,
              stop_rule=False, snapshots=False):
    """Per-step records of the loop.

    Each step solves the unit-load problem (u_entire with Ut = unit,
    crack_propagation.cpp:70), takes Ue from it at t == 0 (:71-80), updates
    the strain history with u_proportion = Ut / unit (:83-85), the Voigt H
    and fd of every level (:86-90), solves the phase field (:91), then
    delta_d_max, d_old = d (:92-93) and the reaction Fw = voxel_size *
    sum_top f_z (:95-96); Ut grows by delta_U and F_U gets
    Fw = u_proportion * FeW (:112-117).  With stop_rule the loop ends after
    the step whose u_proportion * FeW < 0.5 (:119-120).

    schedule: "reference" (Ut = Ue, delta_U = 1.5e-3 Ue), "ue-multiple"
    (Ut = delta_U = du * Ue) or "fixed" (Ut = delta_U = du)."""
    L = obj.maxlevel
    f_norm, fd_norm = 0.0, 0.0
    Ut, delta_U, u_prop = 0.0, 0.0, 1.0
    recs = []
    for t in range(steps):
        f_norm, cyc_u = obj.u_entire(unit, t, f_norm)
        if t == 0:
            if schedule == "reference":
                Ue = obj.U_e()
                Ut = Ue
                delta_U = 1.5e-3 * Ue
            elif schedule == "ue-multiple":
                # fixed increments of du * Ue: the strain history exceeds the
                # threshold from step 0 and the phase field grows
                Ue = obj.U_e()
                Ut, delta_U = du * Ue, du * Ue
            else:
                Ut, delta_U = du, du
            u_prop = Ut / unit
        obj.strain_history(u_prop)
        obj.voigt_all_level(O.H)
        for l in range(L, -1, -1):
            obj.fd_h(l)
        fd_norm, cyc_d = obj.phase_field(t, fd_norm)
        ddmax = obj.delta_d_max()
        dmax = obj.d_max()
        obj.save_d_old()
        FeW = voxel_size * obj.externe_force_sum()
        Ut = Ut + delta_U
        u_prop = Ut / unit
        rec = dict(t=t, cyc_u=cyc_u, cyc_d=cyc_d, f_norm=f_norm, fd_norm=fd_norm, Ut=Ut,
                   Fw=u_prop * FeW, ddmax=ddmax, dmax=dmax, u_prop_solve=(Ut - delta_U) / unit)
        if snapshots:
            rec["u"] = get_vec(obj, O.U, L) / unit
            rec["fu"] = get_vec(obj, O.FU, L) / unit
            rec["d"] = get_field(obj, O.D, L).copy()
            rec["H"] = get_field(obj, O.H, L).copy()
        recs.append(rec)
        if log:
            log(f"  t={t} cyc_u={cyc_u} cyc_d={cyc_d} Fw={u_prop * FeW:.6e} dmax={dmax:.4f}")
        if stop_rule and u_prop * FeW < 0.5:
            break
    return recs


def get_vec(obj, f, l):
    return np.asarray(obj.vec(f, l) if hasattr(obj, "vec") else obj.get(f, l), dtype=np.float64).copy()


def get_field(obj, f, l):
    return np.asarray(obj.field(f, l) if hasattr(obj, "field") else obj.get(f, l), dtype=np.float64)


def rotation_mode(dims, domain):
    """Flattened (3 x nodes) zero-energy mode of the reference's boundary
    conditions: the rigid rotation about the vertical axis through
    (x_b, y_a), which satisfies u_x = 0 at (x_mid, y_a), u_y = 0 at
    (x_b, y_mid) and the u_z conditions (boundary_conditions.cpp:43-58)."""
    nx, ny, nz = dims
    xa, xb, ya, yb = domain[0], domain[1], domain[2], domain[3]
    x = xa + np.arange(nx) * (xb - xa) / (nx - 1)
    y = ya + np.arange(ny) * (yb - ya) / (ny - 1)
    zz, yy, xx = np.meshgrid(np.zeros(nz), y, x, indexing="ij")
    return np.stack([-(yy - ya), xx - xb, np.zeros_like(xx)]).reshape(-1)


def rel_l2_mod_rotation(a, b, dims, domain):
    """Relative L2 distance of two displacement fields after removing the
    component of their difference along the zero-energy rotation mode (the
    mechanical system is singular along it, so its amount is set by
    rounding, not by the problem)."""
    m = rotation_mode(dims, domain)
    e = (np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)).reshape(-1)
    e = e - (e @ m) / (m @ m) * m
    nb = np.linalg.norm(np.asarray(b).reshape(-1))
    return float(np.linalg.norm(e) / (nb if nb > 0 else 1.0))
