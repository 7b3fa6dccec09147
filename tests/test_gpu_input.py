"""The input pipeline on the device (hpcr_set_image: uint8 image -> image_MP
-> K, G, gc -> coarse_VR) against the host restatement hpcr_image_mp plus
the device coarse_VR, and against the oracle's image_MP and coarse_VR.
Bit-identical: both sides run the same FP64 operations in the same order."""
import numpy as np
import pytest

import oracle as O
from gpu_common import require_gpu, say

pytestmark = pytest.mark.gpu


def test_set_image_matches_host_image_mp_and_oracle():
    HA = require_gpu()
    from hpcrack_amd import io as hio
    import hpcrack_amd.workloads as W
    w = W.scaled(W.C2, 33)
    img, _ = w.image()
    assert set(np.unique(img)) == {0, 1, 2}
    tab = w.table()
    say("[gpu] set_image 33^3 three-phase")
    s = HA.Solver(w.maxlevel, w.element_nb, w.domain, **w.solver_kwargs())
    hio.set_image(s, img, tab, w.voxel_size)
    K, G, gc = hio.image_mp(tab, img, w.voxel_size)
    s2 = HA.Solver(w.maxlevel, w.element_nb, w.domain, **w.solver_kwargs())
    W.upload(s2, w, K, G, gc)
    o = O.Oracle(w.maxlevel, w.element_nb, w.domain, **w.solver_kwargs())
    o.image_mp(img, list(W.GRAY), list(tab.E[:3]), list(W.NU), list(W.GC), w.voxel_size)
    for f in (O.K, O.G, O.GC):
        o.coarse_vr_nodal(f)
    L = w.maxlevel
    for l in range(L + 1):
        for f in (O.K, O.G, O.GC):
            a, b = s.get(f, l), s2.get(f, l)
            assert np.array_equal(a, b), (l, f)
            ref = o.field(f, l)
            if l == L:
                # image_MP: the same FP64 operations in the same order
                assert np.array_equal(a, ref), (l, f)
            else:
                # coarse_VR: 27-point means and square roots, summed in another order
                assert np.max(np.abs(a - ref) / np.abs(ref)) < 1e-14, (l, f)


def test_set_image_rejects_unknown_gray_level():
    HA = require_gpu()
    from hpcrack_amd import io as hio
    s = HA.Solver(1, (2, 2, 2), (0, 4, 0, 4, 0, 4))
    img = np.zeros((5, 5, 5), dtype=np.uint8)
    img[2, 2, 2] = 7
    tab = hio.material_table([0, 1], [3e9, 70e9], [0.3, 0.2], [100.0, 1000.0])
    with pytest.raises(hio.IoError):
        hio.set_image(s, img, tab, 1e-6)
