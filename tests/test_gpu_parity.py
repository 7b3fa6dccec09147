"""GPU parity: the HIP path (through the C-ABI) against the oracle on the same
seeded inputs.  Operators and transfers must agree to ~1e-12 relative;
smoother iterates to 1e-11; full solves to 1e-9 relative L2 with identical
cycle counts (+-1 allowed by the acceptance bar, exact expected here)."""
import numpy as np
import pytest

import oracle as O
from gpu_common import make_pair, max_rel, rel_l2, require_gpu, say

pytestmark = pytest.mark.gpu

CASES = [
    # maxlevel, element_nb, domain
    (1, (4, 4, 4), (0.0, 8.0, 0.0, 8.0, 0.0, 8.0)),          # 9^3
    (2, (4, 4, 4), (0.0, 16.0, 0.0, 16.0, 0.0, 16.0)),       # 17^3
    (2, (2, 3, 2), (-1.0, 3.0, 0.0, 6.0, 0.5, 4.5)),        # 9x13x9, anisotropic offsets
    (3, (4, 4, 4), (-16.0, 16.0, -16.0, 16.0, -16.0, 16.0)),  # 33^3
]
IDS = ["9c", "17c", "9x13x9", "33c"]


@pytest.fixture(scope="module")
def HA():
    return require_gpu()


def random_state(o, s, L, seed, d_zero_frac=0.2):
    rng = np.random.default_rng(seed)
    dims = o.dims(L)
    d = rng.uniform(-0.05, 0.9, dims[::-1])
    d[rng.random(dims[::-1]) < d_zero_frac] = 0.0
    fl = rng.integers(0, 3, dims[::-1]).astype(np.float64)
    for f, a in ((O.D, d), (O.FLAG, fl)):
        o.set(f, L, a)
        s.set(f, L, a)
    return d, fl


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_internal_force_and_diagonal(HA, case):
    ml, enb, dom = case
    say(f"[gpu] internal_force {case}")
    o, s, _ = make_pair(HA, ml, enb, dom, seed=1)
    L = ml
    random_state(o, s, L, 2)
    rng = np.random.default_rng(3)
    for l in range(L + 1):
        dims = o.dims(l)
        if l < L:
            Ke = rng.uniform(1e3, 5e4, dims[::-1])
            Ge = rng.uniform(1e3, 3e4, dims[::-1])
            for f, a in ((O.KE, Ke), (O.GE, Ge)):
                o.set(f, l, a)
                s.set(f, l, a)
        x = rng.standard_normal((3,) + dims[::-1])
        o.set(O.P, l, x)
        s.set(O.P, l, x)
        o.internal_force(l, O.P, O.W)
        s.internal_force(l, O.P, O.W)
        assert max_rel(s.get(O.W, l), o.vec(O.W, l)) < 1e-12, f"apply level {l}"
        o.prem_u(l, O.W)
        s.prem_u(l, O.W)
        assert max_rel(s.get(O.W, l), o.vec(O.W, l)) < 1e-13, f"diag level {l}"


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_phase_field_operator_diag_rhs(HA, case):
    ml, enb, dom = case
    say(f"[gpu] fld {case}")
    o, s, _ = make_pair(HA, ml, enb, dom, seed=4, lc=2.0)
    rng = np.random.default_rng(5)
    for l in range(ml + 1):
        dims = o.dims(l)
        Hn = rng.uniform(0, 50, dims[::-1])
        x = rng.standard_normal(dims[::-1])
        for f, a in ((O.H, Hn), (O.PD, x)):
            o.set(f, l, a)
            s.set(f, l, a)
        o.fld(l, O.PD, O.WD)
        s.fld(l, O.PD, O.WD)
        assert max_rel(s.get(O.WD, l), o.field(O.WD, l)) < 1e-12
        o.precondition_m(l, O.WD)
        s.precondition_m(l, O.WD)
        assert max_rel(s.get(O.WD, l), o.field(O.WD, l)) < 1e-13
        o.fd_h(l)
        s.fd_h(l)
        assert max_rel(s.get(O.FD, l), o.field(O.FD, l)) < 1e-13


@pytest.mark.parametrize("case", CASES[:3], ids=IDS[:3])
def test_transfers_and_coarse_coefficients(HA, case):
    ml, enb, dom = case
    say(f"[gpu] transfers {case}")
    o, s, _ = make_pair(HA, ml, enb, dom, seed=6)
    for l in range(ml + 1):
        for f in (O.K, O.G, O.GC):
            assert max_rel(s.get(f, l), o.field(f, l)) < 1e-14, f"coarse_VR field {f} level {l}"
    rng = np.random.default_rng(7)
    for lf in range(1, ml + 1):
        dc, df = o.dims(lf - 1), o.dims(lf)
        rf = rng.standard_normal((3,) + df[::-1])
        fic = rng.standard_normal((3,) + dc[::-1])
        for obj in (o, s):
            obj.set(O.RU, lf, rf)
            obj.set(O.W, lf - 1, fic)
            obj.restrict(lf, O.RU, O.W, O.FU, comp=1)
        assert max_rel(s.get(O.FU, lf - 1, comp=1), o.field(O.FU, lf - 1, 1)) < 1e-14
        uf = rng.standard_normal((3,) + df[::-1])
        uc = rng.standard_normal((3,) + dc[::-1])
        ucold = rng.standard_normal((3,) + dc[::-1])
        for obj in (o, s):
            obj.set(O.U, lf, uf)
            obj.set(O.U, lf - 1, uc)
            obj.set(O.UOLD, lf - 1, ucold)
            obj.prolong(lf, O.U, O.U, O.UOLD, comp=2)
            obj.injection(lf - 1, O.RU, comp=0)
        assert max_rel(s.get(O.U, lf, comp=2), o.field(O.U, lf, 2)) < 1e-14
        assert np.array_equal(s.get(O.RU, lf - 1, comp=0), o.field(O.RU, lf - 1, 0))
    # element coarsening
    df = o.dims(ml)
    Hf = rng.uniform(0.1, 9.0, df[::-1])
    for obj in (o, s):
        obj.set(O.KE, ml, Hf)
        obj.vr_all_level(O.KE)
        obj.set(O.H, ml, Hf)
        obj.voigt_all_level(O.H)
    for l in range(ml):
        assert max_rel(s.get(O.KE, l), o.field(O.KE, l)) < 1e-14
        assert max_rel(s.get(O.H, l), o.field(O.H, l)) < 1e-14


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_split_flag_means_history_elastic_limit(HA, case):
    ml, enb, dom = case
    say(f"[gpu] gauss-point updates {case}")
    o, s, _ = make_pair(HA, ml, enb, dom, seed=8, lc=2.0)
    L = ml
    rng = np.random.default_rng(9)
    dims = o.dims(L)
    u = rng.standard_normal((3,) + dims[::-1]) * 0.01
    d = rng.uniform(0, 0.8, dims[::-1])
    d[rng.random(dims[::-1]) < 0.2] = 0.0
    for obj in (o, s):
        obj.set(O.U, L, u)
        obj.set(O.D, L, d)
    o.mp_choice()
    o.mp_all_element()
    s.mp_choice()
    flo, fls = o.field(O.FLAG, L)[:-1, :-1, :-1], s.get(O.FLAG, L)[:-1, :-1, :-1]
    mism = np.count_nonzero(flo != fls)
    assert mism <= max(1, flo.size // 10000), f"{mism} split flags differ"
    for f in (O.KE, O.GE):
        a, b = s.get(f, L)[:-1, :-1, :-1], o.field(f, L)[:-1, :-1, :-1]
        same = flo == fls
        assert max_rel(a[same], b[same]) < 1e-13
    H0 = rng.uniform(-1e4, 1e3, dims[::-1])
    for obj in (o, s):
        obj.set(O.H, L, H0)
        obj.strain_history(3.7)
    assert max_rel(s.get(O.H, L), o.field(O.H, L)) < 1e-12
    assert s.U_e() == pytest.approx(o.U_e(), rel=1e-12)


@pytest.mark.parametrize("case", CASES[:3], ids=IDS[:3])
@pytest.mark.parametrize("nb", [1, 4])
def test_pcg_smoothers(HA, case, nb):
    ml, enb, dom = case
    say(f"[gpu] pcg nb={nb} {case}")
    o, s, _ = make_pair(HA, ml, enb, dom, seed=10, lc=2.0)
    L = ml
    random_state(o, s, L, 11, d_zero_frac=0.0)
    rng = np.random.default_rng(12)
    dims = o.dims(L)
    u0 = rng.standard_normal((3,) + dims[::-1]) * 0.1
    f0 = rng.standard_normal((3,) + dims[::-1]) * 1e3
    for obj in (o, s):
        obj.set(O.U, L, u0)
        obj.set(O.FU, L, f0)
    fo = o.pcg_u(L, nb, 0.3, 0, 0.0)
    fs = s.pcg_u(L, nb, 0.3, 0, 0.0)
    assert fs == pytest.approx(fo, rel=1e-12)
    assert rel_l2(s.get(O.U, L), o.vec(O.U, L)) < 1e-11
    assert rel_l2(s.get(O.RU, L), o.vec(O.RU, L)) < 1e-10
    assert rel_l2(s.get(O.FU, L), o.vec(O.FU, L)) < 1e-12
    Hn = rng.uniform(0, 20, dims[::-1])
    d0 = rng.uniform(0, 0.5, dims[::-1])
    fd = rng.standard_normal(dims[::-1])
    for obj in (o, s):
        obj.set(O.H, L, Hn)
        obj.set(O.D, L, d0)
        obj.set(O.FD, L, fd)
    fo = o.pcg_d(L, nb)
    fs = s.pcg_d(L, nb)
    assert fs == pytest.approx(fo, rel=1e-13)
    assert rel_l2(s.get(O.D, L), o.field(O.D, L)) < 1e-12
    assert rel_l2(s.get(O.RD, L), o.field(O.RD, L)) < 1e-10


# Coarse levels of at most 512 nodes and 256 elements run a whole smoother
# call in one workgroup; level 0 of every case here takes that path (5^3,
# 3x5x4 ... nodes), level 1 of the 9^3 case does not (512 elements).
SMALL_CASES = [CASES[0], CASES[2], (2, (2, 4, 3), (-1.0, 3.0, 0.0, 8.0, 0.5, 6.5))]


@pytest.mark.parametrize("case", SMALL_CASES, ids=["9c", "9x13x9", "9x17x13"])
@pytest.mark.parametrize("nb", [0, 1, 4, 12])
def test_small_level_smoothers(HA, case, nb):
    ml, enb, dom = case
    say(f"[gpu] small-level pcg nb={nb} {case}")
    o, s, _ = make_pair(HA, ml, enb, dom, seed=20, lc=2.0)
    rng = np.random.default_rng(21)
    for l in (0, 1):
        dims = o.dims(l)
        Ke = rng.uniform(1e3, 5e4, dims[::-1])
        Ge = rng.uniform(1e3, 3e4, dims[::-1])
        u0 = rng.standard_normal((3,) + dims[::-1]) * 0.1
        f0 = rng.standard_normal((3,) + dims[::-1]) * 1e3
        Hn = rng.uniform(0, 20, dims[::-1])
        d0 = rng.uniform(0, 0.5, dims[::-1])
        fd = rng.standard_normal(dims[::-1])
        for obj in (o, s):
            for f, a in ((O.KE, Ke), (O.GE, Ge), (O.U, u0), (O.FU, f0), (O.H, Hn), (O.D, d0), (O.FD, fd)):
                obj.set(f, l, a)
        fo = o.pcg_u(l, nb, 0.3, 0, 0.0)
        fs = s.pcg_u(l, nb, 0.3, 0, 0.0)
        assert fs == fo
        # random right-hand sides are not orthogonal to the zero-energy rotation
        # mode, so long runs lose digits; 12 iterations stay far from that
        tol = 1e-11 if nb <= 4 else 1e-10
        assert rel_l2(s.get(O.U, l), o.vec(O.U, l)) < tol, f"u level {l}"
        if nb <= 4:
            assert rel_l2(s.get(O.RU, l), o.vec(O.RU, l)) < 1e-10, f"r level {l}"
        fo = o.pcg_d(l, nb)
        fs = s.pcg_d(l, nb)
        assert fs == pytest.approx(fo, rel=1e-13)
        assert rel_l2(s.get(O.D, l), o.field(O.D, l)) < tol, f"d level {l}"
        if nb <= 4:
            assert rel_l2(s.get(O.RD, l), o.field(O.RD, l)) < 1e-10, f"R level {l}"


# The x and y pins of the reference's boundary conditions exist only where
# the element counts in x and y are even on every level; the 9x13x9 case has
# 3 elements in y on level 0 and is replaced by one with valid pins.
CYCLE_CASES = [CASES[0], CASES[1], (2, (2, 4, 3), (-1.0, 3.0, 0.0, 8.0, 0.5, 6.5))]
CYCLE_IDS = ["9c", "17c", "9x17x13"]


@pytest.mark.parametrize("case", CYCLE_CASES, ids=CYCLE_IDS)
def test_cycles(HA, case):
    ml, enb, dom = case
    say(f"[gpu] cycles {case}")
    o, s, _ = make_pair(HA, ml, enb, dom, seed=13, lc=2.0)
    L = ml
    random_state(o, s, L, 14, d_zero_frac=0.1)
    rng = np.random.default_rng(15)
    dims = o.dims(L)
    # From a rough random start the reference cycle is chaotic (its coarsest
    # 50-iteration CG over-converges on the singular rotation mode; see
    # test_oracle_multigrid.test_cycle_from_rough_start_is_rounding_sensitive),
    # so the cycle is compared from the physical start u = 0.
    u0 = np.zeros((3,) + dims[::-1])
    for obj in (o, s):
        obj.set(O.U, L, u0)
        obj.set(O.FU, L, np.zeros((3,) + dims[::-1]))
        obj.mech_setup()
    # element range only: the last node layer of element grids is unused
    assert np.array_equal(s.get(O.FLAG, L)[:-1, :-1, :-1], o.field(O.FLAG, L)[:-1, :-1, :-1])
    fo = o.cycle_u(L, 1.0, 0, 0.0)
    fs = s.cycle_u(L, 1.0, 0, 0.0)
    # the coarsest smoother's 50 CG iterations run far past convergence and
    # amplify rounding (observed 4e-9 on f_norm after one cycle)
    assert fs == pytest.approx(fo, rel=1e-7)
    assert rel_l2(s.get(O.U, L), o.vec(O.U, L)) < 1e-9
    Hn = rng.uniform(0, 20, dims[::-1])
    for obj in (o, s):
        obj.set(O.H, L, Hn)
        obj.voigt_all_level(O.H)
        obj.fd_h(L)
        obj.set(O.D, L, np.zeros(dims[::-1]))
    fo = o.cycle_d(L)
    fs = s.cycle_d(L)
    assert fs == pytest.approx(fo, rel=1e-13)
    assert rel_l2(s.get(O.D, L), o.field(O.D, L)) < 1e-10


def test_pcg_smoother_on_a_long_grid(HA):
    """12289 x 5 x 5 nodes: nx + ny + nz exceeds what pass B stages in LDS, so
    its Dirichlet masks come from the axis flags in global memory (the other
    branch of the same pass)."""
    ml, enb, dom = 1, (6144, 2, 2), (0.0, 12288.0, 0.0, 4.0, 0.0, 4.0)
    say(f"[gpu] pcg on a long grid {enb}")
    o, s, _ = make_pair(HA, ml, enb, dom, seed=14, lab_incl=0, lc=2.0)
    L = ml
    assert sum(s.dims(L)) > 12288
    random_state(o, s, L, 15, d_zero_frac=0.0)
    rng = np.random.default_rng(16)
    dims = o.dims(L)
    u0 = rng.standard_normal((3,) + dims[::-1]) * 0.1
    f0 = rng.standard_normal((3,) + dims[::-1]) * 1e3
    for obj in (o, s):
        obj.set(O.U, L, u0)
        obj.set(O.FU, L, f0)
    fo = o.pcg_u(L, 3, 0.3, 0, 0.0)
    fs = s.pcg_u(L, 3, 0.3, 0, 0.0)
    assert fs == pytest.approx(fo, rel=1e-12)
    assert rel_l2(s.get(O.U, L), o.vec(O.U, L)) < 1e-11
    assert rel_l2(s.get(O.RU, L), o.vec(O.RU, L)) < 1e-10


@pytest.mark.parametrize("case", [
    (1, (64, 64, 4), (0.0, 128.0, 0.0, 128.0, 0.0, 8.0)),   # 129x129x9: x - 1 and y - 1 multiples of 64 and 4
    (1, (80, 65, 4), (0.0, 160.0, 0.0, 130.0, 0.0, 8.0)),   # 161x131x9: partial last tiles in x and y
], ids=["129x129x9", "161x131x9"])
def test_pcg_smoother_on_split_tiles(HA, case):
    """Finest levels of 64 or more tiles per plane run pass A on split-output
    tiles (64x4 elements, side parts added up by pass B).  Both edge shapes:
    the last node column and row past every tile's own region (129), and
    last tiles that reach past the grid (161 x 131)."""
    ml, enb, dom = case
    say(f"[gpu] pcg on split tiles {enb}")
    o, s, _ = make_pair(HA, ml, enb, dom, seed=17, lc=2.0)
    L = ml
    random_state(o, s, L, 18, d_zero_frac=0.0)
    rng = np.random.default_rng(19)
    dims = o.dims(L)
    u0 = rng.standard_normal((3,) + dims[::-1]) * 0.1
    f0 = rng.standard_normal((3,) + dims[::-1]) * 1e3
    for obj in (o, s):
        obj.set(O.U, L, u0)
        obj.set(O.FU, L, f0)
    fo = o.pcg_u(L, 4, 0.3, 0, 0.0)
    fs = s.pcg_u(L, 4, 0.3, 0, 0.0)
    assert fs == pytest.approx(fo, rel=1e-12)
    assert rel_l2(s.get(O.U, L), o.vec(O.U, L)) < 1e-11
    assert rel_l2(s.get(O.RU, L), o.vec(O.RU, L)) < 1e-10
    assert rel_l2(s.get(O.FU, L), o.vec(O.FU, L)) < 1e-12


def test_pcg_smoother_on_a_large_coarse_level(HA):
    """A coarse level of 64 or more tiles per plane (129x129x3 under a
    257x257x5 finest level, whose pass A runs on split-output tiles): its
    smoother, with per-element moduli K_e, G_e, on the halo tiles and their
    round-sized z-chunks."""
    ml, enb, dom = 2, (64, 64, 1), (0.0, 256.0, 0.0, 256.0, 0.0, 4.0)
    say(f"[gpu] pcg on a large coarse level {enb}")
    o, s, _ = make_pair(HA, ml, enb, dom, seed=22, lc=2.0)
    rng = np.random.default_rng(23)
    l = 1
    dims = o.dims(l)
    assert dims == (129, 129, 3)
    Ke = rng.uniform(1e3, 5e4, dims[::-1])
    Ge = rng.uniform(1e3, 3e4, dims[::-1])
    u0 = rng.standard_normal((3,) + dims[::-1]) * 0.1
    f0 = rng.standard_normal((3,) + dims[::-1]) * 1e3
    for obj in (o, s):
        for f, a in ((O.KE, Ke), (O.GE, Ge), (O.U, u0), (O.FU, f0)):
            obj.set(f, l, a)
    fo = o.pcg_u(l, 4, 0.3, 0, 0.0)
    fs = s.pcg_u(l, 4, 0.3, 0, 0.0)
    assert fs == fo
    assert rel_l2(s.get(O.U, l), o.vec(O.U, l)) < 1e-11
    assert rel_l2(s.get(O.RU, l), o.vec(O.RU, l)) < 1e-10
