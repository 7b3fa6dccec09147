"""Fused multi-problem cycles (hpcr_fuser): several independent problems,
solved concurrently from their own threads with their multigrid cycles run as
multi-problem launches, against the same problems solved one by one without
a fuser.  Every problem keeps its own grid, reductions and device scalars, so
the results must be identical bit for bit: u, d, the norms, the cycle counts
and the residual histories.  The mix holds problems with the same level
structure (they share launches: 17^3, 17x17x33, 33x17x17 and a second 17^3)
and problems that must not be grouped with them (33^3 has another level count;
17^3 over a 9^3 coarsest level runs its coarsest smoother as tile passes;
129x129x17 runs its finest pass A on split-output tiles, whose w parts pass B
adds up).
In the mixed-precision form every outer iteration of the flexible CG (FP32
V-cycle, FCG passes A and B) goes through the fuser as one unit."""
import threading

import numpy as np
import pytest

import oracle as O
from gpu_common import material_from_labels, require_gpu, say, spheres_labels

pytestmark = pytest.mark.gpu

KW = dict(lc=2.0, small_k=1e-6, tol_u=1e-16, mg_u=(50, 4, 2, 1, 30), mg_d=(10, 2, 1, 1, 30),
          tol_d_rel=1e-16, tol_d_abs=1e-30)

# (maxlevel, element_nb, seed)
PROBLEMS = [
    (2, (4, 4, 4), 60),
    (2, (4, 4, 8), 61),
    (2, (8, 4, 4), 62),
    (2, (4, 4, 4), 63),
    (3, (4, 4, 4), 64),
    (1, (8, 8, 8), 65),
    (2, (32, 32, 4), 66),
]


def build(HA, ml, enb, seed):
    # one voxel per finest element
    dom = (0.0, float(enb[0] * 2 ** ml), 0.0, float(enb[1] * 2 ** ml), 0.0, float(enb[2] * 2 ** ml))
    s = HA.Solver(ml, enb, dom, **KW)
    dims = s.dims(ml)
    lab = spheres_labels(dims, 3, 1.5, max(2.0, min(dims) / 5.0), seed)
    K, G, gc = material_from_labels(lab)
    for f, a in ((O.K, K), (O.G, G), (O.GC, gc)):
        s.set(f, ml, a)
    for f in (O.K, O.G, O.GC):
        s.coarse_vr(f)
    return s


def solve(s, ml, mixed, steps=2):
    """Two staggered increments: unit-load u_entire (or its mixed-precision
    form), Ue, strain history at 3 Ue, phase_field; the second step re-solves
    from the previous state."""
    out = []
    f_norm = 0.0
    fd_norm = 0.0
    for t in range(steps):
        if mixed:
            f_norm, cu = s.u_entire_mixed(1.0, t, f_norm, max_iter=200)
        else:
            f_norm, cu = s.u_entire(1.0, t, f_norm)
        hu = s.history().tolist()
        iu = s.last_stats.smoother_iters
        Ue = s.U_e()
        s.strain_history(3.0 * (t + 1) * Ue)
        s.voigt_all_level(O.H)
        for l in range(ml, -1, -1):
            s.fd_h(l)
        fd_norm, cd = s.phase_field(t, fd_norm)
        hd = s.history().tolist()
        out.append(dict(f_norm=f_norm, cu=cu, hu=hu, iu=iu, Ue=Ue, fd_norm=fd_norm, cd=cd, hd=hd))
    return out, s.get(O.U, ml), s.get(O.D, ml), s.get(O.H, ml)


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
@pytest.mark.parametrize("graphs", [True, False], ids=["graphs", "eager"])
def test_fused_solves_identical_to_solves_alone(graphs, mixed):
    HA = require_gpu()
    say(f"[gpu] fused multi-problem cycles, graphs={graphs}, mixed={mixed}, {len(PROBLEMS)} problems")
    alone = []
    for ml, enb, seed in PROBLEMS:
        s = build(HA, ml, enb, seed)
        alone.append(solve(s, ml, mixed))
        s.close()
    fz = HA.Fuser(window_us=20000, graphs=graphs)
    solvers = [build(HA, ml, enb, seed) for ml, enb, seed in PROBLEMS]
    for s in solvers:
        s.set_fuser(fz)
    fused = [None] * len(PROBLEMS)
    errors = []

    def run(i):
        try:
            fused[i] = solve(solvers[i], PROBLEMS[i][0], mixed)
        except Exception as e:  # reported below
            errors.append((i, repr(e)))

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(PROBLEMS))]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in th), "a fused solve did not return"
    assert not errors, errors
    st = fz.stats()
    say(f"  fuser stats {st}")
    for s in solvers:
        s.close()
    fz.close()
    for i, (a, b) in enumerate(zip(alone, fused)):
        say(f"  problem {i} {PROBLEMS[i][:2]}: cycles u {[x['cu'] for x in a[0]]} d {[x['cd'] for x in a[0]]}")
        assert a[0] == b[0], (i, a[0], b[0])
        for x, y in zip(a[1:], b[1:]):
            assert np.array_equal(x, y), i
    # only problems with a partner of the same structure and size class go
    # through the fuser (33^3 and the 9^3-coarsest 17^3 have none)
    total = sum(x["cu"] + x["cd"] for r in alone for x in r[0])
    solo = sum(x["cu"] + x["cd"] for i in (4, 5) for x in alone[i][0])
    assert 0 < st["cycles"] <= total - solo
    # problems of the same level structure shared launches
    assert st["max_problems"] >= 2
    assert st["problem_launches"] > st["launches"]
    if graphs:
        assert st["graph_replays"] > 0
    else:
        assert st["graph_captures"] == 0 and st["graph_replays"] == 0


def test_fuser_lifecycle():
    HA = require_gpu()
    fz = HA.Fuser()
    s = build(HA, 2, (4, 4, 4), 66)
    s.set_fuser(fz)
    with pytest.raises(HA.HpcrError):
        fz.close()  # a context is still attached
    s.set_fuser(None)
    f_norm, cyc = s.u_entire(1.0, 0, 0.0)  # detached: runs unfused
    assert cyc > 0 and fz.stats()["cycles"] == 0
    s.close()
    fz.close()
