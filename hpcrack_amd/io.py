"""Host-side input formats over the C-ABI of include/hpcrack_io.h:
config.txt, material CSV, uint8 .raw images, geometry, synthetic images,
load-step history files and the load-loop state."""
import ctypes

import numpy as np

from . import capi


class Config(ctypes.Structure):
    _fields_ = [("D", ctypes.c_int), ("np", ctypes.c_int * 3), ("X_start", ctypes.c_double * 3),
                ("IMGname", ctypes.c_char * 1024), ("voxel_nb", ctypes.c_int * 3),
                ("roi_start", ctypes.c_int * 3), ("voxel_size", ctypes.c_double),
                ("element_nb", ctypes.c_int * 3), ("mg_u", capi.MG), ("mg_d", capi.MG), ("lc", ctypes.c_double),
                ("outK", ctypes.c_int), ("outG", ctypes.c_int), ("outgc", ctypes.c_int),
                ("outdesplacement", ctypes.c_int), ("outH", ctypes.c_int), ("outTdata", ctypes.c_int),
                ("Tdatafrequence", ctypes.c_int), ("seen", ctypes.c_ulonglong)]


KEYS = ["D", "np", "X_start", "IMGname", "voxel_nb_X", "voxel_nb_Y", "voxel_nb_Z", "ROI_start_X",
        "ROI_start_Y", "ROI_start_Z", "voxel_size", "element_nb", "ncy_u", "nu0_u", "nu1_u", "nu2_u", "gamma_u",
        "ncy_d", "nu0_d", "nu1_d", "nu2_d", "gamma_d", "lc", "outK", "outG", "outgc", "outdesplacement", "outH",
        "outTdata", "Tdatafrequence"]

MAX_PHASES = 256


class MaterialTable(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("gray", ctypes.c_int * MAX_PHASES), ("E", ctypes.c_double * MAX_PHASES),
                ("nu", ctypes.c_double * MAX_PHASES), ("gc", ctypes.c_double * MAX_PHASES)]


class SynthSpec(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int * 3), ("seed", ctypes.c_ulonglong), ("matrix_label", ctypes.c_int),
                ("n_incl", ctypes.c_int), ("incl_rmin", ctypes.c_double), ("incl_rmax", ctypes.c_double),
                ("incl_label", ctypes.c_int), ("pore_fraction", ctypes.c_double), ("pore_rmin", ctypes.c_double),
                ("pore_rmax", ctypes.c_double), ("pore_label", ctypes.c_int), ("gap", ctypes.c_int),
                ("max_tries", ctypes.c_int)]


class StepState(ctypes.Structure):
    _fields_ = [("t", ctypes.c_int), ("Ut", ctypes.c_double), ("delta_U", ctypes.c_double),
                ("u_prop", ctypes.c_double), ("f_norm", ctypes.c_double), ("Ue", ctypes.c_double)]


_proto_done = False


def _lib():
    global _proto_done
    L = capi.load()
    if not _proto_done:
        dp = ctypes.POINTER(ctypes.c_double)
        ub = ctypes.POINTER(ctypes.c_ubyte)
        L.hpcr_read_config.restype = ctypes.c_int
        L.hpcr_read_config.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(Config)]
        L.hpcr_config_geometry.restype = ctypes.c_int
        L.hpcr_config_geometry.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(capi.Geometry), dp]
        L.hpcr_read_material_csv.restype = ctypes.c_int
        L.hpcr_read_material_csv.argtypes = [ctypes.c_char_p, ctypes.POINTER(MaterialTable)]
        L.hpcr_read_raw.restype = ctypes.c_int
        L.hpcr_read_raw.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_int, ub]
        L.hpcr_write_raw.restype = ctypes.c_int
        L.hpcr_write_raw.argtypes = [ctypes.c_char_p, ub, ctypes.c_size_t]
        L.hpcr_image_mp.restype = ctypes.c_int
        L.hpcr_image_mp.argtypes = [ctypes.POINTER(MaterialTable), ub, ctypes.c_size_t, ctypes.c_double, dp, dp,
                                    dp]
        for fn in (L.hpcr_write_vector_3d, L.hpcr_read_vector_3d):
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, dp, ctypes.c_size_t]
        L.hpcr_write_step_state.restype = ctypes.c_int
        L.hpcr_write_step_state.argtypes = [ctypes.c_char_p, ctypes.POINTER(StepState)]
        L.hpcr_read_step_state.restype = ctypes.c_int
        L.hpcr_read_step_state.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(StepState)]
        for fn in (L.hpcr_set_image, L.hpcr_set_image_noalloc):
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_void_p, ub, ctypes.POINTER(MaterialTable), ctypes.c_double]
        L.hpcr_synth_image.restype = ctypes.c_int
        L.hpcr_synth_image.argtypes = [ctypes.POINTER(SynthSpec), ub, dp]
        _proto_done = True
    return L


class IoError(RuntimeError):
    pass


def read_config(path, strict=True):
    c = Config()
    rc = _lib().hpcr_read_config(path.encode(), 1 if strict else 0, ctypes.byref(c))
    if rc != 0:
        raise IoError(f"cannot read config {path} ({rc})")
    return c


def config_seen(c):
    return {k for i, k in enumerate(KEYS) if c.seen >> i & 1}


def config_geometry(c):
    g = capi.Geometry()
    lc = ctypes.c_double(0.0)
    rc = _lib().hpcr_config_geometry(ctypes.byref(c), ctypes.byref(g), ctypes.byref(lc))
    if rc != 0:
        raise IoError("bad geometry in config")
    return g, lc.value


def read_material_csv(path):
    t = MaterialTable()
    rc = _lib().hpcr_read_material_csv(path.encode(), ctypes.byref(t))
    if rc != 0:
        raise IoError(f"cannot read material table {path} ({rc})")
    return t


def material_table(gray, E, nu, gc):
    t = MaterialTable()
    t.n = len(gray)
    for i in range(t.n):
        t.gray[i], t.E[i], t.nu[i], t.gc[i] = int(gray[i]), float(E[i]), float(nu[i]), float(gc[i])
    return t


def read_raw(path, voxels, roi, dims):
    nx, ny, nz = dims
    out = np.empty((nz, ny, nx), dtype=np.uint8)
    r = (ctypes.c_int * 3)(*roi)
    rc = _lib().hpcr_read_raw(path.encode(), voxels[0], voxels[1], voxels[2], r, nx, ny, nz,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    if rc != 0:
        raise IoError(f"cannot read image {path} ({rc})")
    return out


def write_raw(path, img):
    img = np.ascontiguousarray(img, dtype=np.uint8)
    rc = _lib().hpcr_write_raw(path.encode(), img.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), img.size)
    if rc != 0:
        raise IoError(f"cannot write {path}")


def image_mp(table, img, voxel_size):
    img = np.ascontiguousarray(img, dtype=np.uint8)
    K = np.empty(img.shape)
    G = np.empty(img.shape)
    gc = np.empty(img.shape)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = _lib().hpcr_image_mp(ctypes.byref(table), img.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), img.size,
                              voxel_size, K.ctypes.data_as(dp), G.ctypes.data_as(dp), gc.ctypes.data_as(dp))
    if rc != 0:
        raise IoError("gray level outside the material table")
    return K, G, gc


def synth_image(dims, seed, n_incl=0, incl_r=(4, 12), incl_label=1, pore_fraction=0.0, pore_r=(2, 5),
                pore_label=2, matrix_label=0, gap=1, max_tries=2000000):
    s = SynthSpec()
    for i in range(3):
        s.dims[i] = int(dims[i])
    s.seed = seed
    s.matrix_label = matrix_label
    s.n_incl = n_incl
    s.incl_rmin, s.incl_rmax = incl_r
    s.incl_label = incl_label
    s.pore_fraction = pore_fraction
    s.pore_rmin, s.pore_rmax = pore_r
    s.pore_label = pore_label
    s.gap = gap
    s.max_tries = max_tries
    img = np.empty((dims[2], dims[1], dims[0]), dtype=np.uint8)
    frac = (ctypes.c_double * 2)()
    rc = _lib().hpcr_synth_image(ctypes.byref(s), img.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), frac)
    if rc != 0:
        raise IoError("synthetic image failed")
    return img, (frac[0], frac[1])


def write_vector_3d(dir, name, t, proc, v):
    """write_vector_3D: raw FP64, x-fastest (v shaped (nz, ny, nx))"""
    v = np.ascontiguousarray(v, dtype=np.float64)
    rc = _lib().hpcr_write_vector_3d(dir.encode(), name.encode(), t, proc,
                                     v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), v.size)
    if rc != 0:
        raise IoError(f"cannot write history {name} t={t} ({rc})")


def read_vector_3d(dir, name, t, proc, shape):
    out = np.empty(shape, dtype=np.float64)
    rc = _lib().hpcr_read_vector_3d(dir.encode(), name.encode(), t, proc,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size)
    if rc != 0:
        raise IoError(f"cannot read history {name} t={t} ({rc})")
    return out


def write_step_state(dir, **kw):
    s = StepState(**kw)
    rc = _lib().hpcr_write_step_state(dir.encode(), ctypes.byref(s))
    if rc != 0:
        raise IoError(f"cannot write step state ({rc})")


def read_step_state(dir, t):
    s = StepState()
    rc = _lib().hpcr_read_step_state(dir.encode(), t, ctypes.byref(s))
    if rc != 0:
        raise IoError(f"cannot read step state t={t} ({rc})")
    return {k: getattr(s, k) for k, _ in StepState._fields_}


def set_image(solver, img, table, voxel_size, noalloc=False):
    """input_total on the device: upload the finest level's uint8 image,
    image_MP into K, G, gc and coarse_VR of the three fields.  noalloc:
    through hpcr_set_image_noalloc (no device allocation; fixed scratch)."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    fn = _lib().hpcr_set_image_noalloc if noalloc else _lib().hpcr_set_image
    rc = fn(solver._h, img.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), ctypes.byref(table), voxel_size)
    if rc != 0:
        raise IoError(f"set_image failed ({rc})")
