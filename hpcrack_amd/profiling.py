"""Summaries of rocprofv3 output databases (kernel trace and PMC counters).

    python -m hpcrack_amd.profiling trace DB [N]
        per kernel name and grid: launches, total and average duration
    python -m hpcrack_amd.profiling traffic FETCH_DB WRITE_DB [--kernel SUBSTR] [--copy-bytes B]
        HBM bytes per launch of the largest-grid dispatch of the kernel whose
        name contains SUBSTR, from FETCH_SIZE and WRITE_SIZE (KiB per dispatch)
        scaled by what the same counters read for the plain device copy of B
        bytes recorded in the same run (rocclr's copyBuffer kernel)

    python -m hpcrack_amd.profiling counters DB [--kernel SUBSTR]
        per-dispatch average of every counter recorded for the largest-grid
        dispatch of the kernel whose name contains SUBSTR

    python -m hpcrack_amd.profiling solve WORKLOAD FORM [SOLVES]
        the program to profile: SOLVES (default 1) mechanical solves of a
        benchmark workload (C1..C4) from u = 0, FORM one of plain (FAS,
        1 cycle), f32 (mixed, FP32 V-cycle), f64 (FP64 V-cycle), mfma (FP32
        V-cycle with the matrix-core element contraction), each to the
        workload's tolerance

The databases are the *_results.db files that rocprofv3 writes with
--output-format rocpd (its default on ROCm 7).
"""
import re
import sqlite3
import sys
from collections import defaultdict


def _short(name):
    n = re.sub(r'hpcr::', '', name)
    return re.sub(r'\(.*', '', n)


def trace_summary(db, limit=25):
    cur = sqlite3.connect(db).cursor()
    tot = cur.execute("select sum(duration)/1e6 from kernels").fetchone()[0]
    n = cur.execute("select count(*) from kernels").fetchone()[0]
    rows = cur.execute("select name, count(*), sum(duration)/1e6, avg(duration)/1e3, grid_x, grid_y, grid_z "
                       "from kernels group by name, grid_x, grid_y, grid_z order by sum(duration) desc").fetchall()
    out = [f"total kernel time {tot:.2f} ms, dispatches {n}"]
    for r in rows[:limit]:
        out.append(f"{_short(r[0])[:55]:55s} n={r[1]:5d} tot={r[2]:8.2f}ms ({100 * r[2] / tot:4.1f}%) "
                   f"avg={r[3]:9.2f}us grid=({r[4]},{r[5]},{r[6]})")
    return "\n".join(out)


def _counter(db, cname):
    """{(kernel, grid): [values per dispatch]} of one counter"""
    cur = sqlite3.connect(db).cursor()
    per = defaultdict(lambda: defaultdict(float))
    for name, grid, c, val, disp in cur.execute(
            "select kernel_name, grid_size, counter_name, value, dispatch_id from counters_collection"):
        if c == cname:
            per[(name, grid)][disp] += val
    return {k: list(v.values()) for k, v in per.items()}


def _pick(vals, sub):
    cands = [(k, v) for k, v in vals.items() if sub in k[0]]
    if not cands:
        raise SystemExit(f"no dispatch of a kernel matching {sub!r}")
    k, v = max(cands, key=lambda kv: kv[0][1])
    return k, sum(v) / len(v), len(v)


def traffic(fetch_db, write_db, kernel, copy_bytes):
    f = _counter(fetch_db, "FETCH_SIZE")
    w = _counter(write_db, "WRITE_SIZE")
    (kn, kg), fk, nf = _pick(f, kernel)
    _, wk, nw = _pick(w, kernel)
    _, fc, _ = _pick(f, "copyBuffer")
    _, wc, _ = _pick(w, "copyBuffer")
    ff = copy_bytes / (fc * 1024.0)
    fw = copy_bytes / (wc * 1024.0)
    rd = fk * 1024.0 * ff
    wr = wk * 1024.0 * fw
    lines = [
        f"kernel: {_short(kn)} grid_size={kg} (dispatches: FETCH {nf}, WRITE {nw})",
        f"FETCH_SIZE_KiB_per_dispatch: {fk:.1f}",
        f"WRITE_SIZE_KiB_per_dispatch: {wk:.1f}",
        f"copy_calibration: {copy_bytes} B copy reads FETCH_SIZE {fc:.1f} KiB (factor {ff:.3f}), "
        f"WRITE_SIZE {wc:.1f} KiB (factor {fw:.3f})",
        f"hbm_read_bytes_per_launch: {rd:.4g}",
        f"hbm_write_bytes_per_launch: {wr:.4g}",
        f"bytes_per_launch: {rd + wr:.6g}",
    ]
    return "\n".join(lines)


def counters(db, kernel):
    cur = sqlite3.connect(db).cursor()
    names = sorted({r[0] for r in cur.execute("select distinct counter_name from counters_collection")})
    lines = []
    for c in names:
        (kn, kg), avg, n = _pick(_counter(db, c), kernel)
        if not lines:
            lines.append(f"kernel: {_short(kn)} grid_size={kg} (dispatches: {n})")
        lines.append(f"{c}: {avg:.6g}")
    return "\n".join(lines)


def solve(workload, form, solves=1):
    import time
    import hpcrack_amd as HA
    import hpcrack_amd.workloads as W
    from hpcrack_amd.capi import FU, U
    w = W.WORKLOADS[workload]
    if form == "plain":
        w.mg_u = (w.mg_u[0], w.mg_u[1], w.mg_u[2], w.mg_u[3], 1)
    img, frac, K, G, gc = w.materials()
    s = HA.Solver(w.maxlevel, w.element_nb, w.domain, **w.solver_kwargs())
    W.upload(s, w, K, G, gc)
    L = w.maxlevel
    kw = {"f32": {}, "f64": {"precond_fp64": True}, "mfma": {"element_mfma": True}}
    for i in range(solves):
        for c in range(3):
            s.fill(U, L, 0.0, c)
            s.fill(FU, L, 0.0, c)
        t = time.perf_counter()
        if form == "plain":
            fn, it = s.u_entire(1.0, 0, 0.0)
        else:
            fn, it = s.u_entire_mixed(1.0, 0, 0.0, tol=w.tol_u, max_iter=200, **kw[form])
        s.synchronize()
        print(f"{w.name} {form}: {it} iterations, rel. residual^2 {s.last_stats.r_norm_final / fn:.3e}, "
              f"{time.perf_counter() - t:.3f} s", flush=True)
    return 0


def main(argv):
    if len(argv) >= 2 and argv[0] == "trace":
        print(trace_summary(argv[1], int(argv[2]) if len(argv) > 2 else 25))
        return 0
    if len(argv) >= 3 and argv[0] == "traffic":
        kernel, copy_bytes = "MechOp<2, true, true>", 512 * 1024 * 1024
        rest = argv[3:]
        while rest:
            if rest[0] == "--kernel":
                kernel = rest[1]
            elif rest[0] == "--copy-bytes":
                copy_bytes = int(rest[1])
            rest = rest[2:]
        print(traffic(argv[1], argv[2], kernel, copy_bytes))
        return 0
    if len(argv) >= 3 and argv[0] == "solve":
        return solve(argv[1], argv[2], int(argv[3]) if len(argv) > 3 else 1)
    if len(argv) >= 2 and argv[0] == "counters":
        kernel = argv[3] if len(argv) >= 4 and argv[2] == "--kernel" else "MechOp<2, true, true>"
        print(counters(argv[1], kernel))
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
