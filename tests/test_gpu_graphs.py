"""Multigrid cycles replayed as HIP graphs (hpcr_set_graphs) against the
same solves launched kernel by kernel: identical results, bit for bit, over a
staggered loop whose load changes every step (new captures) with eager work
and host synchronisation between the replays."""
import numpy as np
import pytest

import oracle as O
from gpu_common import make_pair, require_gpu, say
from staggered import staggered

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager_solves():
    HA = require_gpu()
    kw = dict(lc=2.0, small_k=1e-6, tol_u=1e-16, mg_u=(50, 4, 2, 1, 30), mg_d=(10, 2, 1, 1, 30),
              tol_d_rel=1e-16, tol_d_abs=1e-30)
    dom = (0.0, 32.0, 0.0, 32.0, 0.0, 32.0)
    say("[gpu] graphs vs eager, 33^3, 3 staggered steps")
    _, eager, _ = make_pair(HA, 3, (4, 4, 4), dom, seed=50, lab_incl=4, **kw)
    _, graph, _ = make_pair(HA, 3, (4, 4, 4), dom, seed=50, lab_incl=4, **kw)
    graph.set_graphs(True)
    re_ = staggered(eager, 3, 1e-6, 2.0, schedule="ue-multiple", du=2.0)
    rg = staggered(graph, 3, 1e-6, 2.0, schedule="ue-multiple", du=2.0)
    for a, b in zip(re_, rg):
        say(f"  t={a['t']} cyc_u {a['cyc_u']}/{b['cyc_u']} cyc_d {a['cyc_d']}/{b['cyc_d']}")
        assert a == b
    for f in (O.U, O.D, O.H, O.FU):
        assert np.array_equal(eager.get(f, 3), graph.get(f, 3))
    # a t = 0 solve again from u = 0: the captured cycle of the first step is replaced
    for s in (eager, graph):
        for c in range(3):
            s.fill(O.U, 3, 0.0, c)
            s.fill(O.FU, 3, 0.0, c)
    fe, ce = eager.u_entire(1.0, 0, 0.0)
    fg, cg = graph.u_entire(1.0, 0, 0.0)
    assert (fe, ce) == (fg, cg)
    assert eager.last_stats.smoother_iters == graph.last_stats.smoother_iters
    assert list(eager.last_stats.history[:ce]) == list(graph.last_stats.history[:cg])
    assert np.array_equal(eager.get(O.U, 3), graph.get(O.U, 3))


def test_reduction_fence_gives_identical_solves():
    """The reductions' release/acquire fence form (hpcr_set_reduction_fence)
    against the default completion-wait form: a staggered step and a mixed
    solve give bit-identical results (the same partials, summed in the same
    fixed order)."""
    HA = require_gpu()
    kw = dict(lc=2.0, small_k=1e-6, tol_u=1e-16, mg_u=(50, 4, 2, 1, 30), mg_d=(10, 2, 1, 1, 30),
              tol_d_rel=1e-16, tol_d_abs=1e-30)
    dom = (0.0, 32.0, 0.0, 32.0, 0.0, 32.0)
    say("[gpu] reduction fence vs completion wait, 33^3")
    _, plain, _ = make_pair(HA, 3, (4, 4, 4), dom, seed=51, lab_incl=4, **kw)
    _, fenced, _ = make_pair(HA, 3, (4, 4, 4), dom, seed=51, lab_incl=4, **kw)
    fenced.set_reduction_fence(True)
    ra = staggered(plain, 2, 1e-6, 2.0, schedule="ue-multiple", du=2.0)
    rb = staggered(fenced, 2, 1e-6, 2.0, schedule="ue-multiple", du=2.0)
    assert ra == rb
    for f in (O.U, O.D, O.H, O.FU):
        assert np.array_equal(plain.get(f, 3), fenced.get(f, 3))
    out = []
    for s in (plain, fenced):
        for c in range(3):
            s.fill(O.U, 3, 0.0, c)
            s.fill(O.FU, 3, 0.0, c)
        out.append((s.u_entire_mixed(1.0, 0, 0.0, tol=1e-16), s.get(O.U, 3)))
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])
