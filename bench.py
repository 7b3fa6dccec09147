#!/usr/bin/env python3
"""Benchmark of the plain FP64 MG-PCG mechanical solve on one MI355X.

Workload (default): the 257^3-node three-phase image (200 stiff inclusions
r ~ U(3,10), pores r ~ U(2,5) to 5 %, seed 1002), element_nb 4 (7 levels),
multigrid u = (nu0 50, nu1 4, nu2 2, V-cycle), k = 1e-6, relative residual
1e-8 (1e-16 in the solver's squared convention), at most 30 cycles.

A step is one u_entire solve from u = 0 (t = 0): split flag, element moduli,
coarse homogenisation, diagonals, initial residual, V-cycles to the tolerance
or the cap, and the Dirichlet reactions.  All inputs are in HBM before the
timed region.  value = 3 N_fine * (V-cycles in the timed steps) / time.

The JSON line also carries the roofline of the dominant kernel (the fused
p-update + degraded-elasticity apply + p.Ap pass on the finest level, timed
with HIP events on the solver's stream) and a CPU baseline: the oracle run on
the host cores over a bounded sample (see cpu_baseline.sample).

Outside the timed region it also measures the metric's other parts:
- `mixed`: the FP64 flexible CG with the FP32 V-cycle preconditioner, timed to
  the same 1e-8 tolerance, against the same flexible CG with the identical
  V-cycle in FP64 (the like-for-like FP32 vs FP64 comparison at one
  tolerance), and against the plain FAS solve at the residual that solve
  reaches (it stalls above the tolerance at its 30-cycle cap);
- `serve`: p50/p99 latency and solves/s of the request server on the
  256-request stream (Poisson 20/s, node dims per axis from {33, 65, 129,
  257}), mixed-precision and FP64 forms.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per fine node of the fused pass A (p = z + beta p_old;
# w = A p; p.w): read z, p_old (2 x 24), K, G, d (24), flag (1); write p, w
# (2 x 24).  Pass B (u, r update, z = r/D) moves 192 B per node; one smoother
# iteration moves 313 B per node in total.
PASS_A_BYTES_PER_NODE = 121.0
PASS_B_BYTES_PER_NODE = 192.0
# smoother residual r = f - A u, z = r/D: read u, f, 1/D (3 x 24), K, G, d, flag
# (25); write r, z (2 x 24)
RESID_BYTES_PER_NODE = 145.0
# canonical bytes of one Jacobi-PCG iteration (pass A + pass B, 313 B per fine
# node) and of one mechanical V-cycle with nu_u = (50, 4, 2): 2144 B per fine
# node on the finest level (6 iterations, 2 residual set-ups, transfers) plus
# 322 on all coarser levels
PCG_ITERATION_BYTES_PER_NODE = 313.0
VCYCLE_BYTES_PER_FINE_NODE = 2466.0
METRIC = "MDOF·iter/s + solves/s, FP32-MG vs FP64 @1e-8 residual; p50/p99 serve latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(path):
    """'bytes_per_launch: X' line of a committed rocprofv3 PMC summary."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.strip().startswith("bytes_per_launch:"):
                try:
                    return float(line.split(":", 1)[1].strip().split()[0])
                except ValueError:
                    return None
    return None


def copy_calibration(mib):
    """One plain device-to-device copy of `mib` MiB (HIP's blit kernel), for
    scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for rc in (hip.hipMalloc(ctypes.byref(a), n), hip.hipMalloc(ctypes.byref(b), n),
               hip.hipMemset(a, 1, n), hip.hipDeviceSynchronize(),
               hip.hipMemcpy(b, a, n, 3), hip.hipDeviceSynchronize(),  # 3: device to device
               hip.hipFree(a), hip.hipFree(b)):
        if rc != 0:
            raise RuntimeError(f"copy calibration: HIP error {rc}")
    log(f"[bench] copy calibration: {mib} MiB device-to-device copy")


def host_cores():
    """Logical CPUs of the host and those this process may run on."""
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    return os.cpu_count() or 1, aff


def cpu_baseline(workload, n_sample, cycles, threads, ref_loops=False):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    import hpcrack_amd.workloads as W

    w = W.scaled(workload, n_sample)
    w.mg_u = (w.mg_u[0], w.mg_u[1], w.mg_u[2], w.mg_u[3], cycles)
    img, frac, K, G, gc = w.materials()
    O.set_threads(threads)
    O.set_ref_loops(ref_loops)
    try:
        o = O.Oracle(w.maxlevel, w.element_nb, w.domain, **w.solver_kwargs())
        L = w.maxlevel
        for f, a in ((O.K, K), (O.G, G), (O.GC, gc)):
            o.set(f, L, a)
        for f in (O.K, O.G, O.GC):
            o.coarse_vr_nodal(f)
        t0 = time.perf_counter()
        _, cyc = o.u_entire(1.0, 0, 0.0)
        dt = time.perf_counter() - t0
    finally:
        O.set_ref_loops(False)
    dof = 3 * int(np.prod(w.dims))
    nproc, aff = host_cores()
    form = ("in the reference's loop form (Gauss-point coefficients re-evaluated 8 times per element, "
            "critical-section scatter-adds)" if ref_loops else
            "in its own loop form, which evaluates each element's Gauss-point coefficients once where the "
            "reference does it 8 times and scatters without critical sections, so it is faster than the "
            "reference it restates")
    return {
        "value": dof * cyc / dt / 1e6,
        "unit": "MDOF·iter/s",
        "cores": O.lib().ora_get_threads(),
        "kind": "port",
        "nproc": nproc,
        "affinity_cpus": aff,
        "sample": f"oracle u_entire (OpenMP, {O.lib().ora_get_threads()} threads) on the same generator at "
                  f"{n_sample}^3 nodes ({w.n_incl} inclusions, {100 * frac[1]:.1f}% pores), t=0 from u=0, "
                  f"{cyc} V-cycles in {dt:.1f} s, {form}",
    }


def serve_stream(mixed, requests, rate, timeout):
    """The request server on the serving stream; its JSON summary line."""
    exe = os.path.join(ROOT, "hpcrack_amd", "hpcrack_serve")
    cmd = [exe, "--requests", str(requests), "--rate", repr(float(rate))] + (["--mixed"] if mixed else ["--fp64"])
    t = time.perf_counter()
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"hpcrack_serve failed ({r.returncode}): {r.stderr[-2000:]}")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    log(f"[bench] serve {'mixed' if mixed else 'fp64'}: {out['solves_per_s']} solves/s, p50 "
        f"{out['latency_ms_p50']} ms, p99 {out['latency_ms_p99']} ms ({time.perf_counter() - t:.1f} s)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="C2")
    ap.add_argument("--ncy", type=int, default=30)
    ap.add_argument("--cpu-sample", type=int, default=257, help="nodes per axis of the CPU sample (0: skip)")
    ap.add_argument("--cpu-cycles", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-ref-sample", type=int, default=65,
                    help="nodes per axis of the second CPU sample, in the reference's loop form (0: skip)")
    ap.add_argument("--cpu-sweep", default="",
                    help="only time CPU samples, one JSON line each, no GPU: comma-separated "
                         "THREADS:CYCLES:FORM:NODES with FORM o (the oracle's loops) or r (the reference's)")
    ap.add_argument("--traffic-file", default="",
                    help="rocprofv3 PMC summary of pass A on this workload "
                         "(default: profiles/traffic_mech_pcg_a_<workload>_<n>.txt; traffic is null without one)")
    ap.add_argument("--mixed", type=int, default=1,
                    help="also time the mixed-precision solve (FP64 flexible CG + FP32 V-cycle) to the same "
                         "tolerance (0: skip)")
    ap.add_argument("--mixed-cheb", type=int, default=0,
                    help="Chebyshev degree of the smoother (0: chosen from the modulus contrast)")
    ap.add_argument("--mixed-ratio", type=float, default=6.0, help="Chebyshev smoothing range lambda_max/lambda_min")
    ap.add_argument("--mixed-coarse-iters", type=int, default=0,
                    help="iterations of the preconditioner's coarsest solve (0: nu0_u)")
    ap.add_argument("--mixed-ref-tol", type=float, default=1e-22,
                    help="squared relative residual of the untimed FP64 reference solve that the mixed and plain "
                         "solutions are compared with (0: skip)")
    ap.add_argument("--serve", type=int, default=1, help="also run the request server on the serving stream (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, one device copy of this size (counter calibration)")
    args = ap.parse_args()

    if args.cpu_sweep:
        import hpcrack_amd.workloads as W
        w = W.WORKLOADS[args.workload]
        for spec in args.cpu_sweep.split(","):
            th, cyc, form, n = spec.split(":")
            t = time.perf_counter()
            r = cpu_baseline(w, int(n), int(cyc), int(th), ref_loops=(form == "r"))
            r["wall_s"] = round(time.perf_counter() - t, 1)
            r["omp_proc_bind"] = os.environ.get("OMP_PROC_BIND", "")
            print(json.dumps(r), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import hpcrack_amd as HA
    import hpcrack_amd.workloads as W
    from hpcrack_amd.capi import FU, K_MECH_PCG_A, K_MECH_PCG_B, K_MECH_RESID, U

    w = W.WORKLOADS[args.workload]
    w.mg_u = (w.mg_u[0], w.mg_u[1], w.mg_u[2], w.mg_u[3], args.ncy)
    t = time.perf_counter()
    img, frac, K, G, gc = w.materials()
    log(f"[bench] {w.name}: dims {w.dims}, levels {w.maxlevel + 1}, inclusions {frac[0]:.4f}, "
        f"pores {frac[1]:.4f} ({time.perf_counter() - t:.1f} s)")
    s = HA.Solver(w.maxlevel, w.element_nb, w.domain, **w.solver_kwargs())
    W.upload(s, w, K, G, gc)
    L = w.maxlevel
    N = int(np.prod(w.dims))
    log(f"[bench] device memory {s.device_bytes() / 2**30:.2f} GiB")

    def zero_state():
        s.fill(U, L, 0.0, 0), s.fill(U, L, 0.0, 1), s.fill(U, L, 0.0, 2)
        s.fill(FU, L, 0.0, 0), s.fill(FU, L, 0.0, 1), s.fill(FU, L, 0.0, 2)

    def step():
        zero_state()
        fn, cyc = s.u_entire(1.0, 0, 0.0)
        return fn, cyc

    for i in range(args.warmup):
        fn, cyc = step()
        log(f"[bench] warmup {i}: {cyc} cycles, rel. residual^2 {s.last_stats.r_norm_final / fn:.3e}")

    s.timing(True)
    s.timing_reset()
    s.synchronize()
    t0 = time.perf_counter()
    cycles = 0
    rels = []
    for i in range(args.steps):
        fn, cyc = step()
        cycles += cyc
        rels.append(s.last_stats.r_norm_final / fn)
    s.synchronize()
    t1 = time.perf_counter()
    s.timing(False)
    elapsed = t1 - t0
    log(f"[bench] timed {args.steps} solves, {cycles} cycles in {elapsed:.3f} s")

    kern = {}
    for name, kid in (("mech_pcg_a", K_MECH_PCG_A), ("mech_pcg_b", K_MECH_PCG_B), ("mech_resid", K_MECH_RESID)):
        ms, n = s.timing_query(kid, L)
        kern[name] = {"total_ms": ms, "launches": n, "avg_ms": ms / n if n else None}
    for name, b in (("mech_pcg_a", PASS_A_BYTES_PER_NODE), ("mech_pcg_b", PASS_B_BYTES_PER_NODE),
                    ("mech_resid", RESID_BYTES_PER_NODE)):
        k = kern[name]
        k["algorithmic_GBs"] = round(b * N / (k["avg_ms"] / 1e3) / 1e9, 1) if k["avg_ms"] else None
    a = kern["mech_pcg_a"]
    avg_s = a["avg_ms"] / 1e3
    achieved = PASS_A_BYTES_PER_NODE * N / avg_s / 1e9
    traffic_file = args.traffic_file or os.path.join(ROOT, "profiles",
                                                     f"traffic_mech_pcg_a_{args.workload}_{w.dims[0]}.txt")
    traffic = read_traffic(traffic_file)
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": "mech_pcg_a (finest level)",
        "avg_launch_ms": round(a["avg_ms"], 4),
        "algorithmic_bytes_per_launch": PASS_A_BYTES_PER_NODE * N,
        "fine_level_share": round(sum(k["total_ms"] for k in kern.values()) / (elapsed * 1e3), 3),
    }
    # the same accounting over a whole PCG iteration (pass A + pass B) and over
    # the whole V-cycle (canonical bytes over the measured time per cycle,
    # set-up of each solve included)
    b = kern["mech_pcg_b"]
    if b["avg_ms"]:
        it_gbs = PCG_ITERATION_BYTES_PER_NODE * N / ((a["avg_ms"] + b["avg_ms"]) / 1e3) / 1e9
        roofline["pcg_iteration"] = {"bytes_per_node": PCG_ITERATION_BYTES_PER_NODE, "achieved": round(it_gbs, 1),
                                     "frac": round(it_gbs / HBM_PEAK_GBS, 4)}
    if cycles:
        vc_gbs = VCYCLE_BYTES_PER_FINE_NODE * N / (elapsed / cycles) / 1e9
        roofline["vcycle"] = {"bytes_per_fine_node": VCYCLE_BYTES_PER_FINE_NODE, "achieved": round(vc_gbs, 1),
                              "frac": round(vc_gbs / HBM_PEAK_GBS, 4), "ms_per_cycle": round(elapsed / cycles * 1e3, 3)}
    value = 3 * N * cycles / elapsed / 1e6

    mixed = None
    if args.mixed:
        u_plain = s.get(U, L)
        plain_rel = max(rels)
        mode = W.rotation_mode(w, L)

        def err(u, ref):
            """relative L2 distance with the zero-energy rotation mode projected out"""
            du = (u - ref).ravel()
            du = du - (du @ mode) / (mode @ mode) * mode
            return float(np.linalg.norm(du) / np.linalg.norm(ref))

        def timed_fcg(tol, p64, label, mfma=False):
            """warmup + timed flexible-CG solves from u = 0: (solves/s, ms, iterations per solve, worst rel)"""
            kw = dict(tol=tol, max_iter=200, cheb_degree=cheb[0], cheb_ratio=cheb[1], cheb_upper=cheb[2],
                      precond_fp64=p64, element_mfma=mfma, coarse_iters=args.mixed_coarse_iters)
            for i in range(args.warmup):
                zero_state()
                fm, it = s.u_entire_mixed(1.0, 0, 0.0, **kw)
                log(f"[bench] {label} warmup {i}: {it} iterations, rel. residual^2 "
                    f"{s.last_stats.r_norm_final / fm:.3e}")
            s.synchronize()
            t0 = time.perf_counter()
            its, worst = 0, 0.0
            for i in range(args.steps):
                zero_state()
                fm, it = s.u_entire_mixed(1.0, 0, 0.0, **kw)
                its += it
                worst = max(worst, s.last_stats.r_norm_final / fm)
            s.synchronize()
            el = time.perf_counter() - t0
            log(f"[bench] {label}: {its} iterations in {el:.3f} s")
            return args.steps / el, el / args.steps * 1e3, its / args.steps, worst

        # the Chebyshev setting every form below uses (the contrast-chosen one by default)
        cheb = s.mixed_resolve(args.mixed_cheb, args.mixed_ratio)
        log(f"[bench] mixed smoother: degree {cheb[0]}, range {cheb[1]}, safety {cheb[2]} (contrast {cheb[3]:.3g})")
        # FP32 V-cycle preconditioner (the mixed-precision form) to the tolerance
        sps32, ms32, it32, rel32 = timed_fcg(w.tol_u, False, "mixed FP32 V-cycle")
        u_mixed = s.get(U, L)
        # the identical V-cycle in FP64, same flexible CG, same tolerance
        sps64, ms64, it64, rel64 = timed_fcg(w.tol_u, True, "FP64 V-cycle")
        u_p64 = s.get(U, L)
        # the FP32 V-cycle with the finest level's element contraction on the matrix cores
        spsm, msm, itm, relm = timed_fcg(w.tol_u, False, "mixed FP32 V-cycle, matrix-core element", mfma=True)
        u_mfma = s.get(U, L)
        # the mixed form timed to the residual the plain FAS solve reaches
        spsq, msq, itq, relq = timed_fcg(plain_rel, False, "mixed FP32 V-cycle to the plain solve's residual")
        # untimed reference: the same FP64 system converged far past the tolerance (the flexible
        # CG's outer operator, residual and updates are FP64; the preconditioner's precision does
        # not change the converged solution)
        ref = None
        if args.mixed_ref_tol > 0:
            zero_state()
            fr, itr = s.u_entire_mixed(1.0, 0, 0.0, tol=args.mixed_ref_tol, max_iter=400,
                                       cheb_degree=cheb[0], cheb_ratio=cheb[1], cheb_upper=cheb[2],
                                       precond_fp64=True)
            ref = {"rel_residual_squared": s.last_stats.r_norm_final / fr, "fcg_iterations": itr,
                   "preconditioner": "f64 V-cycle"}
            u_ref = s.get(U, L)
            log(f"[bench] FP64 reference: {itr} iterations, rel. residual^2 {ref['rel_residual_squared']:.3e}")
        plain_sps = args.steps / elapsed
        mixed = {
            "solves_per_s": round(sps32, 3),
            "ms_per_solve": round(ms32, 2),
            "fcg_iterations_per_solve": it32,
            "mdof_iter_s": round(3 * N * it32 * sps32 / 1e6, 2),
            "final_rel_residual_squared": rel32,
            "reached_tol": bool(rel32 <= w.tol_u),
            "u_rel_l2_vs_fp64_converged": err(u_mixed, u_ref) if ref else None,
            # like for like: the same flexible CG and V-cycle with the preconditioner in FP64,
            # to the same tolerance
            "fp64_vcycle": {
                "solves_per_s": round(sps64, 3),
                "ms_per_solve": round(ms64, 2),
                "fcg_iterations_per_solve": it64,
                "final_rel_residual_squared": rel64,
                "reached_tol": bool(rel64 <= w.tol_u),
                "u_rel_l2_vs_fp64_converged": err(u_p64, u_ref) if ref else None,
            },
            "speedup_fp32_vs_fp64_vcycle_same_tol": round(sps32 / sps64, 3),
            # the same FP32 V-cycle with the finest level's element contraction as batched FP32
            # matrix-core products (v_mfma_f32_16x16x4_f32), FP32 accumulation
            "mfma_element": {
                "solves_per_s": round(spsm, 3),
                "ms_per_solve": round(msm, 2),
                "fcg_iterations_per_solve": itm,
                "final_rel_residual_squared": relm,
                "reached_tol": bool(relm <= w.tol_u),
                "u_rel_l2_vs_fp64_converged": err(u_mfma, u_ref) if ref else None,
            },
            # against the plain FAS solve, which stops at its cycle cap above the tolerance: the
            # mixed form timed to the residual the plain solve reaches, and the plain solve's state
            "plain_fp64_final_rel_residual_squared": plain_rel,
            "plain_fp64_reached_tol": bool(plain_rel <= w.tol_u),
            "plain_u_rel_l2_vs_fp64_converged": err(u_plain, u_ref) if ref else None,
            "at_plain_residual": {
                "tol": plain_rel,
                "solves_per_s": round(spsq, 3),
                "fcg_iterations_per_solve": itq,
                "speedup_vs_plain": round(spsq / plain_sps, 3),
            },
            "solves_ratio_vs_plain_capped_unconverged": round(sps32 / plain_sps, 3),
            "fp64_converged_reference": ref,
            "cheb_degree": cheb[0],
            "cheb_ratio": cheb[1],
            "cheb_upper": cheb[2],
            "cheb_chosen_by": "modulus contrast %.3g" % cheb[3] if args.mixed_cheb == 0 else "argument",
            "coarse_iters": args.mixed_coarse_iters or w.mg_u[0],
            "dtype": "f32 V-cycle (Chebyshev-Jacobi, transfers, operators), f64 outer flexible CG, f64 coarsest",
        }
        log(f"[bench] mixed: {sps32:.3f} solves/s; FP64 V-cycle {sps64:.3f} solves/s "
            f"(x{sps32 / sps64:.2f}); at the plain residual {spsq:.3f} solves/s")

    cpu = None
    if args.cpu_sample > 0:
        log(f"[bench] cpu baseline: oracle at {args.cpu_sample}^3, {args.cpu_cycles} cycles, "
            f"{args.cpu_threads} threads")
        cpu = cpu_baseline(w, args.cpu_sample, args.cpu_cycles, args.cpu_threads)
        log(f"[bench] cpu baseline {cpu['value']:.3f} MDOF-iter/s")
        if args.cpu_ref_sample > 0:
            # the same algorithm with the reference's own loop costs, on a smaller sample
            r = cpu_baseline(w, args.cpu_ref_sample, 1, args.cpu_threads, ref_loops=True)
            cpu["reference_loop_form"] = {"value": r["value"], "sample": r["sample"]}
            log(f"[bench] cpu baseline, reference loop form: {r['value']:.3f} MDOF-iter/s")

    serve = None
    if args.serve:
        # the stream at its arrival rate (latency) and as one burst (capacity: every request
        # waiting at t = 0, so solves/s is the server's throughput, not the arrival rate)
        serve = {"stream": f"{args.serve_requests} requests, Poisson {args.serve_rate}/s, node dims per axis from "
                           "{33,65,129,257}, inclusion fraction U(0.10,0.30), one load step each, 8 slots, "
                           "HBM budget 200 GB",
                 "mixed": serve_stream(True, args.serve_requests, args.serve_rate, 600),
                 "mixed_capacity": serve_stream(True, args.serve_requests, 1e5, 600),
                 "fp64": serve_stream(False, args.serve_requests, args.serve_rate, 600),
                 "fp64_capacity": serve_stream(False, args.serve_requests, 1e5, 600)}

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "MDOF·iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: plain FP64 u_entire, {w.dims[0]}x{w.dims[1]}x{w.dims[2]} nodes, "
                        f"{w.maxlevel + 1} levels, nu_u=({w.mg_u[0]},{w.mg_u[1]},{w.mg_u[2]}) V-cycle",
            "dof": 3 * N,
            "cycles_per_solve": cycles / args.steps,
            "solves_per_s": round(args.steps / elapsed, 3),
            "tol_u_squared": w.tol_u,
            "ncy": args.ncy,
            "final_rel_residual_squared": max(rels),
            "inclusion_fraction": round(frac[0], 4),
            "pore_fraction": round(frac[1], 4),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "serve": serve,
        "kernels_finest_level": kern,
    }
    print(json.dumps(out), flush=True)
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)


if __name__ == "__main__":
    main()
