"""Pins of the staggered load-step loop (crack_propagation.cpp:65-121) as
run over the oracle, and the measured rounding floor of the damaged
schedule that bounds the GPU comparisons.

The loop restatement lives in tests/staggered.py.  Its trajectory is pinned
by what the reference's loop fixes whatever the solver: the load sequence
Ut (:71-80, :112-113), the reaction Fw written to F_U (:95-96, :116-117) as
an independent assembly of A u on the top face, the monotone strain history
(strain_history.cpp:197-207), delta_d_max and d_max against the d
snapshots (phase_field_max_search.cpp:17-53), the stop rule (:119-120),
and the phase-field and unit-load solutions against the assembled systems.
"""
import numpy as np
import pytest

import oracle as O
import textbook as T
from gpu_common import material_from_labels, rel_l2, spheres_labels
from staggered import rel_l2_mod_rotation, staggered

DOM9 = (0.0, 8.0, 0.0, 8.0, 0.0, 8.0)
DOM17 = (0.0, 16.0, 0.0, 16.0, 0.0, 16.0)
KW = dict(lc=2.0, small_k=1e-6, tol_u=1e-16, mg_u=(50, 4, 2, 1, 30), mg_d=(10, 2, 1, 1, 30),
          tol_d_rel=1e-16, tol_d_abs=1e-30)


def two_phase(ml, dom, seed, n_incl=3):
    o = O.Oracle(ml, (4, 4, 4), dom, **KW)
    dims = o.dims(ml)
    lab = spheres_labels(dims, n_incl, 1.5, max(2.0, min(dims) / 5.0), seed)
    K, G, gc = material_from_labels(lab)
    for f, a in ((O.K, K), (O.G, G), (O.GC, gc)):
        o.set(f, ml, a)
    for f in (O.K, O.G, O.GC):
        o.coarse_vr_nodal(f)
    return o, K, G, gc


def test_trajectory_closed_forms_9_cubed():
    """Reference schedule and a damaging 2 Ue schedule at 9^3."""
    for schedule, du in (("reference", None), ("ue-multiple", 2.0)):
        o, K, G, gc = two_phase(1, DOM9, 5)
        L = 1
        dims = o.dims(L)
        steps = 4
        recs = []
        flags, d_before = [], []
        # run step by step to keep the split flags and d each u solve used
        f_norm = fd_norm = 0.0
        Ut = delta = 0.0
        d_prev = np.zeros(dims[::-1])
        H_prev = np.zeros(dims[::-1])
        for t in range(steps):
            d_before.append(o.field(O.D, L).copy())
            f_norm, _ = o.u_entire(1.0, t, f_norm)
            flags.append(o.field(O.FLAG, L).copy())
            r_final = o.history("u")[-1]
            if t == 0:
                Ue = Ue0 = o.U_e()
                Ut, delta = (Ue, 1.5e-3 * Ue) if schedule == "reference" else (du * Ue, du * Ue)
            u_prop = Ut
            o.strain_history(u_prop)
            H = o.field(O.H, L).copy()
            assert np.all(H >= H_prev)  # H_e = max(H_e, ...): monotone
            H_prev = H
            o.voigt_all_level(O.H)
            for l in range(L, -1, -1):
                o.fd_h(l)
            fd_norm, _ = o.phase_field(t, fd_norm)
            d = o.field(O.D, L).copy()
            assert o.delta_d_max() == np.max(d - d_prev)
            assert o.d_max() == np.max(d)
            o.save_d_old()
            d_prev = d
            FeW = 1e-6 * o.externe_force_sum()
            Ut += delta
            recs.append(dict(t=t, Ut=Ut, Fw=Ut * FeW, u=o.vec(O.U, L).copy(), d=d, H=H, r_final=r_final,
                             f_norm=f_norm))
        # the same loop through the shared restatement gives the same records
        o2, *_ = two_phase(1, DOM9, 5)
        rr = staggered(o2, steps, 1e-6, 2.0, schedule=schedule, du=du)
        for a, b in zip(recs, rr):
            assert a["Ut"] == b["Ut"] and a["Fw"] == b["Fw"]
        # load sequence
        for t, r in enumerate(recs):
            exp = Ue0 * (1 + 1.5e-3 * (t + 1)) if schedule == "reference" else du * Ue0 * (t + 2)
            assert r["Ut"] == pytest.approx(exp, rel=1e-14)
        # Fw: voxel_size * u_proportion (after the increment) * sum of the
        # reactions f_z = (A u)_z on the top face, A assembled independently
        # with the split flags and d of that step's u solve
        h = o.h(L)
        nx, ny, nz = dims
        for t, r in enumerate(recs):
            coef = T.fine_coefficients(dims, h, K, G, d_before[t], flags[t], 1e-6)
            A = T.elastic_matrix(dims, h, coef)
            Au = A @ r["u"].reshape(-1)
            Nn = nx * ny * nz
            top = [2 * Nn + i + nx * (j + ny * (nz - 1)) for j in range(ny) for i in range(nx)]
            assert r["Fw"] == pytest.approx(r["Ut"] * 1e-6 * np.sum(Au[top]), rel=1e-9)
            # u solves the unit-load system: the true free-DOF residual of the
            # independent assembly is the one the solve reported (the FAS
            # cycle may stop at its cap above the tolerance)
            vals = T.dirichlet(dims, h, (0.0, 0.0, 0.0), (8.0, 8.0, 8.0), 1.0)
            free = np.setdiff1d(np.arange(3 * Nn), np.array(sorted(vals)))
            rr = float(np.sum(Au[free] ** 2))
            assert rr == pytest.approx(r["r_final"], rel=1e-3, abs=1e-22 * r["f_norm"])
            assert rr <= 1e-9 * r["f_norm"]
            # d solves the phase-field system of H (Voigt-free on the finest level)
            Ad = T.phase_matrix(dims, h, r["H"], gc, 2.0)
            fd = np.zeros(dims[::-1])
            He = r["H"][:-1, :-1, :-1] * 2.0 * (h[0] * h[1] * h[2]) / 8.0
            for a in (0, 1):
                for b in (0, 1):
                    for c in (0, 1):
                        fd[a:a + nz - 1, b:b + ny - 1, c:c + nx - 1] += He
            res = fd.reshape(-1) - Ad @ r["d"].reshape(-1)
            if np.linalg.norm(fd) > 0:
                assert np.linalg.norm(res) <= 1e-7 * np.linalg.norm(fd)
            else:
                assert np.all(r["d"] == 0)
        if schedule == "ue-multiple":
            assert recs[-1]["d"].max() > 0.5


def test_stop_rule_ends_the_loop():
    """The loop ends after the first step whose F_U force u_prop * FeW is
    below 0.5 (crack_propagation.cpp:119-120).  The voxel size scales Fw, so
    at 5 um the 2 Ue schedule at 17^3 crosses 0.5 at step 2 as the crack
    grows (Fw 0.77, 1.1, 0.40 N)."""
    o, *_ = two_phase(2, DOM17, 21)
    recs = staggered(o, 6, 5e-6, 2.0, schedule="ue-multiple", du=2.0, stop_rule=True)
    fw = [r["Fw"] for r in recs]
    assert all(f >= 0.5 for f in fw[:-1]) and fw[-1] < 0.5
    assert len(recs) == 3


def floor_runs(steps=4, du=2.0):
    """The 2 Ue schedule at 17^3 run twice by the oracle, with the unit load
    1 and 1 + 2^-52 (an exact-arithmetic change of 1e-16): the per-step
    distance of u (modulo the zero-energy rotation), d and H between the
    runs is the rounding floor of this schedule."""
    runs = []
    for unit in (1.0, np.nextafter(1.0, 2.0)):
        o, *_ = two_phase(2, DOM17, 21)
        runs.append(staggered(o, steps, 1e-6, 2.0, schedule="ue-multiple", du=du, unit=unit, snapshots=True))
    dims = (17, 17, 17)
    out = []
    for a, b in zip(*runs):
        out.append(dict(u_raw=rel_l2(b["u"], a["u"]), u=rel_l2_mod_rotation(b["u"], a["u"], dims, DOM17),
                        d=rel_l2(b["d"], a["d"]), H=rel_l2(b["H"], a["H"]),
                        crack=int((a["d"] > 0.95).sum()), cyc=(a["cyc_u"], b["cyc_u"], a["cyc_d"], b["cyc_d"])))
    return out


@pytest.mark.slow
def test_rounding_floor_of_the_damaged_schedule():
    """Measured floor (values at the time of writing, 17^3, seed 21):
    step  u (mod rotation)  u raw    d        H        crack set d > 0.95
    0     4e-13             5.6e-10  1e-14    7e-15    0
    1     1e-13             5.6e-10  2e-15    5e-15    0
    2     5e-12             4.8e-10  4e-14    3e-13    123 voxels
    3     2.3e-8            9.6e-8   2.9e-11  2.8e-10  1003 voxels
    Raw u moves by 5e-10 at every step along the rotation mode, which the
    system leaves free; modulo that mode the floor is 1e-13 until the
    specimen is nearly cut through (step 3: reaction down 10x)."""
    fl = floor_runs()
    for t, f in enumerate(fl):
        print(t, f)
    assert all(f["cyc"][0] == f["cyc"][1] and f["cyc"][2] == f["cyc"][3] for f in fl)
    assert fl[0]["u_raw"] > 1e-11  # the free rotation takes the rounding
    assert max(f["u"] for f in fl[:3]) < 1e-10
    assert fl[3]["u"] < 1e-6
    assert max(f["d"] for f in fl) < 1e-9 and max(f["H"] for f in fl) < 1e-8
    # the crack set is non-empty from step 2 and not the whole specimen
    assert fl[1]["crack"] == 0 and 0 < fl[2]["crack"] < fl[3]["crack"] < 17 ** 3
