"""Pins the oracle's Gauss-point energy routines, split flag and material input
with closed forms.

Reference routines pinned: strain_history, U_e, Phi_positive, Phi_negative, W,
Phi_c (strain_history.cpp:22-210); MP_choice_element
(material_selection.cpp:19-45); image_MP (image_input.cpp:120-142);
findPosition / material_property (csv_reader.cpp:16-82); externe_force_sum
(mechanical_right_hnd.cpp:40-59); delta_d_max, d_max
(phase_field_max_search.cpp:17-53).
"""
import numpy as np
import pytest

import oracle as O

DOM = (0.0, 2.0, 0.0, 2.0, 0.0, 2.0)


def linear_u(o, l, E):
    nx, ny, nz = o.dims(l)
    hx, hy, hz = o.h(l)
    zz, yy, xx = np.meshgrid(np.arange(nz) * hz, np.arange(ny) * hy, np.arange(nx) * hx, indexing="ij")
    X = np.stack([xx, yy, zz])
    return np.einsum("cm,mzyx->czyx", E, X)


def phis(K, G, S):
    tr = np.trace(S)
    dev = S - tr / 3 * np.eye(3)
    ddp = np.sum(dev * dev)
    pp = 0.5 * K * tr * tr + G * ddp if tr >= 0 else G * ddp
    pn = 0.0 if tr >= 0 else 0.5 * K * tr * tr
    return tr, pp, pn


STRAINS = {
    "tension": np.array([[0.01, 0.0, 0.0], [0.0, 0.002, 0.0], [0.0, 0.0, 0.03]]),
    "compression_shear": np.array([[-0.02, 0.001, 0.0], [0.001, -0.01, 0.0], [0.0, 0.0, -0.01]]),
    "mild_compression": np.array([[-0.001, 0.02, 0.01], [0.02, 0.0005, 0.0], [0.01, 0.0, 0.0002]]),
}


@pytest.mark.parametrize("name", list(STRAINS))
def test_split_flag_for_uniform_strain(name):
    E = STRAINS[name]
    o = O.Oracle(0, (2, 2, 2), DOM)
    dims = o.dims(0)
    Kc, Gc = 2.0, 0.9
    o.set(O.K, 0, np.full(dims[::-1], Kc))
    o.set(O.G, 0, np.full(dims[::-1], Gc))
    o.set(O.U, 0, linear_u(o, 0, E))
    o.mp_choice()
    tr, pp, pn = phis(Kc, Gc, 0.5 * (E + E.T))
    exp = 0 if tr >= 0 else 1
    if pp < pn:
        exp = 2
    assert np.all(o.field(O.FLAG, 0)[:-1, :-1, :-1] == exp)
    # the three cases are all exercised by the parametrisation
    assert name != "compression_shear" or exp == 2
    assert name != "mild_compression" or exp == 1


@pytest.mark.parametrize("name", list(STRAINS))
def test_strain_history_and_elastic_limit_closed_form(name):
    E = STRAINS[name]
    S = 0.5 * (E + E.T)
    lc = 1.5
    o = O.Oracle(0, (2, 2, 2), DOM, lc=lc)
    dims = o.dims(0)
    Kc, Gc, gcv = 2.0, 0.9, 0.3
    o.set(O.K, 0, np.full(dims[::-1], Kc))
    o.set(O.G, 0, np.full(dims[::-1], Gc))
    o.set(O.GC, 0, np.full(dims[::-1], gcv))
    o.set(O.U, 0, linear_u(o, 0, E))
    uprop = 2.5
    H0 = np.full(dims[::-1], -1e9)
    o.set(O.H, 0, H0)
    o.strain_history(uprop)
    _, pp, _ = phis(Kc, Gc, uprop * S)
    exp = pp - 0.5 * gcv / lc
    assert np.allclose(o.field(O.H, 0)[:-1, :-1, :-1], exp, rtol=1e-12)
    # monotone: a smaller load never lowers H
    o.strain_history(0.1)
    assert np.allclose(o.field(O.H, 0)[:-1, :-1, :-1], exp, rtol=1e-12)
    # elastic limit: sqrt(phi_c / W) with W = sigma:eps / 2
    tr = np.trace(S)
    sig = Kc * tr * np.eye(3) + 2 * Gc * (S - tr / 3 * np.eye(3))
    W = 0.5 * np.sum(sig * S)
    assert o.U_e() == pytest.approx(np.sqrt(0.5 * gcv / lc / W), rel=1e-12)


def test_elastic_limit_is_min_over_gauss_points_and_skips_nan():
    o = O.Oracle(0, (2, 2, 2), DOM, lc=1.0)
    dims = o.dims(0)
    o.set(O.K, 0, np.full(dims[::-1], 2.0))
    o.set(O.G, 0, np.full(dims[::-1], 1.0))
    o.set(O.GC, 0, np.full(dims[::-1], 1.0))
    # zero displacement: W = 0 -> sqrt(phi_c/0) = inf, never below 1e10 start
    assert o.U_e() == 1e10


def test_image_mp_stepwise_interpolation_of_shipped_table():
    # data/Material_Property.csv of the reference: gray 0, 90, 91, 255
    gray = [0, 90, 91, 255]
    E = [21e9, 21e9, 210e9, 210e9]
    nu = [0.2, 0.2, 0.3, 0.3]
    gc = [180.0, 180.0, 1730.0, 1730.0]
    o = O.Oracle(0, (1, 1, 1), (0, 1, 0, 1, 0, 1))
    img = np.array([0, 45, 90, 91, 150, 255, 254, 1], dtype=np.uint8).reshape(2, 2, 2)
    vs = 2.53e-6
    o.image_mp(img, gray, E, nu, gc, vs)
    # integer division: only the upper breakpoint itself takes the next value
    def expect(x):
        pos = next(i for i in range(3) if gray[i] <= x <= gray[i + 1])
        fac = (x - gray[pos]) // (gray[pos + 1] - gray[pos])
        e = E[pos] + fac * (E[pos + 1] - E[pos])
        v = nu[pos] + fac * (nu[pos + 1] - nu[pos])
        g = gc[pos] + fac * (gc[pos + 1] - gc[pos])
        return vs * e / (3 * (1 - 2 * v)), vs * e / (2 * (1 + v)), g
    K, G, GC = o.field(O.K, 0), o.field(O.G, 0), o.field(O.GC, 0)
    for q, x in enumerate(img.reshape(-1)):
        k, g, c = expect(int(x))
        assert K.reshape(-1)[q] == pytest.approx(k, rel=1e-15)
        assert G.reshape(-1)[q] == pytest.approx(g, rel=1e-15)
        assert GC.reshape(-1)[q] == pytest.approx(c, rel=1e-15)
    # gray 45 is inside [0, 90]: stepwise, so it keeps the gray-0 values
    assert GC.reshape(-1)[1] == 180.0 and GC.reshape(-1)[4] == 1730.0


def test_reaction_sum_and_max_searches():
    o = O.Oracle(0, (2, 2, 2), DOM)
    dims = o.dims(0)
    rng = np.random.default_rng(2)
    fz = rng.standard_normal(dims[::-1])
    f = np.zeros((3,) + dims[::-1])
    f[2] = fz
    o.set(O.FU, 0, f)
    assert o.externe_force_sum() == pytest.approx(fz[-1].sum(), rel=1e-14)
    d = rng.uniform(-0.2, 0.8, dims[::-1])
    o.set(O.D, 0, d)
    assert o.d_max() == d.max()
    assert o.delta_d_max() == max(0.0, d.max())
    o.save_d_old()
    o.set(O.D, 0, d - 0.1)
    assert o.delta_d_max() == 0.0
