"""hpcr_set_image_noalloc (the allocation-free image set-up) against
hpcr_set_image, through the C-ABI: K, G and gc on every level are
bit-identical (image_MP and coarse_VR are integer-table and FP64 work in the
same order on both paths).

Two calls in a row on one device, the second with a smaller image: the first
image (257 x 129 x 129 nodes, 4.3 MB) streams through the 4 MiB staging area
in two chunks, the second (65 x 33 x 33) in one, so the second call runs on
scratch that still holds the first call's contents.  An image with a gray
level outside the table is rejected after a good call and before another."""
import numpy as np
import pytest

from gpu_common import require_gpu, say

pytestmark = pytest.mark.gpu


def _case(W, n, n_incl, seed):
    """Three-phase image of n x (n+1)/2 x (n+1)/2 nodes (element_nb 8, 4, 4)."""
    w = W.Workload(f"noalloc-{n}", n, seed, n_incl=n_incl, incl_r=(2.0, 6.0), pore_fraction=0.03,
                   pore_r=(1.5, 3.0), element_nb=(8, 4, 4))
    img, _ = w.image()
    assert img.shape == tuple(reversed(w.dims)) and w.dims[0] == n
    return w, img


def _solvers(HA, w):
    kw = w.solver_kwargs()
    return (HA.Solver(w.maxlevel, w.element_nb, w.domain, **kw), HA.Solver(w.maxlevel, w.element_nb, w.domain, **kw))


def test_set_image_noalloc_matches_set_image_twice():
    HA = require_gpu()
    import hpcrack_amd.workloads as W
    from hpcrack_amd import io as hio
    import oracle as O
    cases = [_case(W, 257, 120, 1), _case(W, 65, 6, 2)]
    assert cases[0][1].size > 4 << 20 > cases[1][1].size
    for k, (w, img) in enumerate(cases):
        say(f"[gpu] set_image_noalloc call {k + 1}: dims {w.dims} ({img.size} bytes)")
        a, b = _solvers(HA, w)
        tab = w.table()
        hio.set_image(a, img, tab, w.voxel_size, noalloc=True)
        hio.set_image(b, img, tab, w.voxel_size)
        assert set(np.unique(img)) == {0, 1, 2}
        for l in range(w.maxlevel + 1):
            for f in (O.K, O.G, O.GC):
                x, y = a.get(f, l), b.get(f, l)
                assert np.array_equal(x, y), (k, l, f)
        assert np.all(a.get(O.K, w.maxlevel) > 0)
        a.close()
        b.close()


def test_set_image_noalloc_rejects_unknown_gray_level_then_recovers():
    HA = require_gpu()
    from hpcrack_amd import io as hio
    import oracle as O
    tab = hio.material_table([0, 1], [3e9, 70e9], [0.3, 0.2], [100.0, 1000.0])
    s = HA.Solver(1, (2, 2, 2), (0, 4, 0, 4, 0, 4))
    good = np.zeros((5, 5, 5), dtype=np.uint8)
    good[1:3, 1:3, 1:3] = 1
    bad = good.copy()
    bad[2, 2, 2] = 7
    say("[gpu] set_image_noalloc: good, bad, good")
    hio.set_image(s, good, tab, 1e-6, noalloc=True)
    k1 = s.get(O.K, 1)
    with pytest.raises(hio.IoError):
        hio.set_image(s, bad, tab, 1e-6, noalloc=True)
    # the flag left set by the rejected call is cleared by the next one
    hio.set_image(s, good, tab, 1e-6, noalloc=True)
    assert np.array_equal(s.get(O.K, 1), k1)
    s.close()
