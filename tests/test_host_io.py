"""Host-side formats and the C-ABI surface (CPU only: no device calls).

Pins: Parameters::readFromFile (configuration.cpp:14-97) on the reference's
own shipped config.txt (tests/golden/reference_config.txt, a data file of the
reference), material_property / findPosition (csv_reader.cpp:16-82) on its
shipped Material_Property.csv, image_MP (image_input.cpp:120-142) against the
oracle's restatement, input (image_input.cpp:33-76) against numpy slicing, the
level stack of main.cpp:52-68 / image_input.cpp:21-29 / initialization.cpp:19-47.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def hio():
    from hpcrack_amd import io
    return io


def test_capi_exports_every_declared_symbol():
    from hpcrack_amd import capi
    L = capi.load()
    names = set()
    for h in ("hpcrack.h", "hpcrack_io.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b(hpcr_\w+)\s*\(", txt))
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # every prototype of the Python binding is declared in the header
    assert set(capi.PROTOTYPES) <= names


def test_no_device_means_clean_error():
    import hpcrack_amd as HA
    if HA.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(HA.HpcrError) as e:
        HA.Solver(1, (2, 2, 2), (0, 2, 0, 2, 0, 2))
    assert e.value.code == -2


def test_fuser_argument_checks():
    """The fusion engine's entry points reject bad arguments before any
    device call (so this runs without a GPU), and with no device visible the
    engine cannot be created."""
    import hpcrack_amd as HA
    from hpcrack_amd import capi
    L = capi.load()
    out = ctypes.c_void_p()
    assert L.hpcr_fuser_create(0, -1, ctypes.byref(out)) == capi.ERR_ARG
    assert L.hpcr_fuser_create(0, 100, None) == capi.ERR_ARG
    assert L.hpcr_fuser_destroy(None) == capi.OK
    assert L.hpcr_fuser_set_graphs(None, 1) == capi.ERR_ARG
    assert L.hpcr_fuser_get_stats(None, ctypes.byref(capi.FuserStats())) == capi.ERR_ARG
    assert L.hpcr_set_fuser(None, None) == capi.ERR_ARG
    if HA.device_count() == 0:
        with pytest.raises(HA.HpcrError) as e:
            HA.Fuser()
        assert e.value.code == capi.ERR_DEVICE


def test_reference_config_strict_drops_spaced_keys(hio):
    c = hio.read_config(os.path.join(GOLD, "reference_config.txt"), strict=True)
    seen = hio.config_seen(c)
    # "lc = 3.;", "outTdata = 1;" and "Tdatafrequence = 5; " are never parsed
    assert "lc" not in seen and "outTdata" not in seen and "Tdatafrequence" not in seen
    assert {"np", "X_start", "IMGname", "voxel_size", "element_nb", "ncy_u", "gamma_d", "outK"} <= seen
    assert list(c.np) == [1, 1, 1]
    assert list(c.X_start) == [-64.0, -64.0, -64.0]
    assert c.IMGname.decode() == "data/inclusion_more_129_129_129.raw"
    assert list(c.voxel_nb) == [129, 129, 129] and list(c.roi_start) == [0, 0, 0]
    assert c.voxel_size == 2.53e-6
    assert list(c.element_nb) == [4, 4, 4]
    assert (c.mg_u.ncy, c.mg_u.nu0, c.mg_u.nu1, c.mg_u.nu2, c.mg_u.gamma) == (12, 50, 4, 2, 1)
    assert (c.mg_d.ncy, c.mg_d.nu0, c.mg_d.nu1, c.mg_d.nu2, c.mg_d.gamma) == (12, 10, 2, 1, 1)
    assert c.outK == 1 and c.outG == 0 and c.lc == 0.0


def test_reference_config_lenient_and_geometry(hio):
    c = hio.read_config(os.path.join(GOLD, "reference_config.txt"), strict=False)
    assert c.lc == 3.0 and c.outTdata == 1 and c.Tdatafrequence == 5
    g, lc = hio.config_geometry(c)
    assert g.maxlevel == 5
    assert (g.xa, g.xb, g.ya, g.yb, g.za, g.zb) == (-64.0, 64.0, -64.0, 64.0, -64.0, 64.0)
    assert list(g.element_nb) == [4, 4, 4]
    assert lc == 3.0  # lc * h_fine, h_fine = 128 / 128


def test_config_geometry_quirks(tmp_path, hio):
    # ROI start shifts the domain end (xb = X_start + ROI + element_nb*2^L)
    p = tmp_path / "c.txt"
    p.write_text("voxel_nb_X=40\nvoxel_nb_Y=40\nvoxel_nb_Z=40\nROI_start_X=3\nX_start=0.5,0,0\n"
                 "element_nb=2,2,2\nlc=2.7\n")
    c = hio.read_config(str(p))
    assert c.lc == 2.0  # stoi
    g, lc = hio.config_geometry(c)
    # nb_ele_f = 40 - 3 - 1 = 36: 4, 8, 16, 32 <= 36 -> maxlevel 4
    assert g.maxlevel == 4
    assert g.xb == 0.5 + 3 + 2 * 16 and g.yb == 32.0
    assert lc == 2.0 * 1.0


def test_material_csv_and_image_mp_match_oracle(hio):
    t = hio.read_material_csv(os.path.join(GOLD, "reference_material_property.csv"))
    assert t.n == 4 and list(t.gray[:4]) == [0, 90, 91, 255]
    assert list(t.E[:4]) == [21e9, 21e9, 210e9, 210e9] and list(t.gc[:4]) == [180.0, 180.0, 1730.0, 1730.0]
    img = np.random.default_rng(0).integers(0, 256, (5, 6, 7), dtype=np.uint8)
    K, G, gc = hio.image_mp(t, img, 2.53e-6)
    o = O.Oracle(0, (6, 5, 4), (0, 6, 0, 5, 0, 4))
    o.image_mp(img, [0, 90, 91, 255], [21e9, 21e9, 210e9, 210e9], [0.2, 0.2, 0.3, 0.3],
               [180.0, 180.0, 1730.0, 1730.0], 2.53e-6)
    assert np.array_equal(K, o.field(O.K, 0)) and np.array_equal(G, o.field(O.G, 0))
    assert np.array_equal(gc, o.field(O.GC, 0))
    with pytest.raises(hio.IoError):
        hio.image_mp(hio.material_table([10, 20], [1, 2], [0.3, 0.3], [1, 1]), img, 1.0)


def test_raw_roi_read(tmp_path, hio):
    full = np.random.default_rng(1).integers(0, 256, (9, 8, 7), dtype=np.uint8)
    p = str(tmp_path / "img.raw")
    hio.write_raw(p, full)
    sub = hio.read_raw(p, (7, 8, 9), (1, 2, 3), (5, 4, 6))
    assert np.array_equal(sub, full[3:9, 2:6, 1:6])
    with pytest.raises(hio.IoError):
        hio.read_raw(p, (7, 8, 10), (0, 0, 0), (7, 8, 10))  # short file


def test_reference_example_image_reads_like_numpy(hio):
    path = "/root/reference/data/inclusion_more_129_129_129.raw"
    if not os.path.exists(path):
        pytest.skip("reference tree absent")
    img = hio.read_raw(path, (129, 129, 129), (0, 0, 0), (129, 129, 129))
    ref = np.fromfile(path, dtype=np.uint8).reshape(129, 129, 129)
    assert np.array_equal(img, ref)
    vals, counts = np.unique(img, return_counts=True)
    assert len(vals) == 9 and counts[vals == 255][0] / img.size > 0.99


def test_synthetic_generator_is_deterministic_and_non_overlapping(hio):
    a, fa = hio.synth_image((40, 36, 33), 7, n_incl=6, incl_r=(3, 5), pore_fraction=0.02, pore_r=(1.5, 2.5))
    b, fb = hio.synth_image((40, 36, 33), 7, n_incl=6, incl_r=(3, 5), pore_fraction=0.02, pore_r=(1.5, 2.5))
    c, _ = hio.synth_image((40, 36, 33), 8, n_incl=6, incl_r=(3, 5), pore_fraction=0.02, pore_r=(1.5, 2.5))
    assert np.array_equal(a, b) and fa == fb and not np.array_equal(a, c)
    assert set(np.unique(a)) <= {0, 1, 2}
    assert fa[1] >= 0.02 and abs(fa[0] - np.mean(a == 1)) < 1e-12
    # labels of different spheres never touch: every 1-voxel is at least the
    # gap away from every 2-voxel
    one = np.argwhere(a == 1)
    two = np.argwhere(a == 2)
    d2 = ((one[:, None, :] - two[None, :, :]) ** 2).sum(-1)
    assert d2.min() >= 2


def test_history_files_match_write_vector_3d_layout(hio, tmp_path):
    # write_vector_3D (write_read_history.cpp:18-52): raw float64, x-fastest,
    # <dir>/<name>_t_<t>_Proc_<id>.dat; a name holding ".dat" is the path itself
    rng = np.random.default_rng(7)
    v = rng.standard_normal((3, 4, 5))  # (nz, ny, nx)
    hio.write_vector_3d(str(tmp_path), "H", 15, 0, v)
    path = tmp_path / "H_t_15_Proc_0.dat"
    raw = path.read_bytes()
    assert len(raw) == v.size * 8
    assert np.array_equal(np.frombuffer(raw, dtype="<f8").reshape(v.shape), v)
    assert np.array_equal(hio.read_vector_3d(str(tmp_path), "H", 15, 0, v.shape), v)
    explicit = str(tmp_path / "u_field.dat")
    hio.write_vector_3d("unused", explicit, 0, 0, v)
    assert open(explicit, "rb").read() == raw
    # a file of the wrong length is refused rather than read partially
    with pytest.raises(hio.IoError):
        hio.read_vector_3d(str(tmp_path), "H", 15, 0, (3, 4, 4))
    with pytest.raises(hio.IoError):
        hio.read_vector_3d(str(tmp_path), "H", 15, 0, (3, 4, 6))
    with pytest.raises(hio.IoError):
        hio.read_vector_3d(str(tmp_path), "d", 15, 0, v.shape)


def test_step_state_round_trip_is_exact(hio, tmp_path):
    st = dict(t=10, Ut=1.2345678901234567e-3, delta_U=1.5e-3 * 0.1234567890123, u_prop=1.2345678901234567e-3,
              f_norm=3.0000000000000004e17, Ue=0.1234567890123)
    hio.write_step_state(str(tmp_path), **st)
    assert hio.read_step_state(str(tmp_path), 10) == st
    with pytest.raises(hio.IoError):
        hio.read_step_state(str(tmp_path), 11)
    (tmp_path / "state_t_12.txt").write_text("t 12\nUt 1\n")
    with pytest.raises(hio.IoError):
        hio.read_step_state(str(tmp_path), 12)


def config_seen_all(hio, c):
    return {"D", "np", "X_start", "IMGname", "voxel_size", "element_nb", "ncy_u", "ncy_d", "lc"} <= hio.config_seen(c)


def test_workload_case_files_read_back_as_the_workload(tmp_path, hio):
    """write_case (python -m hpcrack_amd.workloads write-case) writes what the
    driver's readers parse back into the workload's geometry, multigrid
    parameters, material table and image."""
    import hpcrack_amd.workloads as W
    w = W.scaled(W.C2, 33, "C2@33")
    img = W.write_case(w, str(tmp_path))
    c = hio.read_config(str(tmp_path / "config.txt"))  # strict: every key parsed
    g, lc = hio.config_geometry(c)
    assert g.maxlevel == w.maxlevel and list(g.element_nb) == list(w.element_nb)
    assert (g.xa, g.xb, g.ya, g.yb, g.za, g.zb) == w.domain
    assert lc == w.lc_voxels
    mg = lambda m: (m.nu0, m.nu1, m.nu2, m.gamma, m.ncy)
    assert mg(c.mg_u) == tuple(w.mg_u) and mg(c.mg_d) == tuple(w.mg_d)
    assert config_seen_all(hio, c)
    t = hio.read_material_csv(str(tmp_path / "mat.csv"))
    ref = w.table()
    assert t.n == ref.n
    for k in range(t.n):
        assert (t.gray[k], t.E[k], t.nu[k], t.gc[k]) == (ref.gray[k], ref.E[k], ref.nu[k], ref.gc[k])
    got = hio.read_raw(str(tmp_path / "img.raw"), w.dims, (0, 0, 0), w.dims)
    assert np.array_equal(got, img)
    K0, G0, gc0 = hio.image_mp(ref, img, w.voxel_size)
    K1, G1, gc1 = hio.image_mp(t, got, c.voxel_size)
    assert np.array_equal(K0, K1) and np.array_equal(G0, G1) and np.array_equal(gc0, gc1)
