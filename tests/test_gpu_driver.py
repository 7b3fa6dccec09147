"""The C++ host driver (hpcrack_amd/hpcrack_run, crack_propagation + main over
the C-ABI) against the oracle's staggered loop on a seeded 17^3 image given
as config.txt + uint8 .raw + material CSV."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from gpu_common import require_gpu, say
from gpu_common import rel_l2
from staggered import rel_l2_mod_rotation, staggered

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "hpcrack_amd", "hpcrack_run")


def parse_fu(path):
    rows = []
    for line in open(path):
        f = dict(kv.split("=") for kv in line.replace(" ", "").split(","))
        rows.append({k: float(v) for k, v in f.items()})
    return rows


def write_case(tmp_path, n=17, seed=31):
    from hpcrack_amd import io as hio
    img, _ = hio.synth_image((n, n, n), seed, n_incl=3, incl_r=(2.0, 4.0))
    hio.write_raw(str(tmp_path / "img.raw"), img)
    (tmp_path / "mat.csv").write_text("Gray level,0,1\nE/Pa,3000000000,70000000000\nnu,0.3,0.2\ngc/(N/m),100,1000\n")
    (tmp_path / "config.txt").write_text(
        "D=3\nnp=1,1,1\nX_start=0,0,0\nIMGname=%s\nvoxel_nb_X=%d\nvoxel_nb_Y=%d\nvoxel_nb_Z=%d\n"
        "ROI_start_X=0\nROI_start_Y=0\nROI_start_Z=0\nvoxel_size=1e-6\nelement_nb=4,4,4\n"
        "ncy_u=30\nnu0_u=50\nnu1_u=4\nnu2_u=2\ngamma_u=1\nncy_d=30\nnu0_d=10\nnu1_d=2\nnu2_d=1\ngamma_d=1\n"
        "lc=2\n" % (tmp_path / "img.raw", n, n, n))
    return img


def driver_args(tmp_path, steps, out="F_U.dat"):
    return [DRIVER, str(tmp_path / "config.txt"), "--csv", str(tmp_path / "mat.csv"), "--steps", str(steps),
            "--small-k", "1e-6", "--tol-u", "1e-16", "--tol-d", "1e-16", "--out", str(tmp_path / out),
            "--stop-force", "-1e300"]


@pytest.mark.parametrize("schedule", ["fixed", "reference", "ue-multiple"])
def test_driver_matches_oracle_staggered_loop(tmp_path, schedule):
    require_gpu()
    img = write_case(tmp_path)
    steps = 4
    args = driver_args(tmp_path, steps)
    if schedule == "fixed":
        args += ["--du", "0.05"]
    elif schedule == "ue-multiple":
        args += ["--du-ue", "2"]
    say(f"[gpu] driver {schedule}")
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = parse_fu(tmp_path / "F_U.dat")
    assert len(got) == steps

    o = O.Oracle(2, (4, 4, 4), (0, 16, 0, 16, 0, 16), small_k=1e-6, lc=2.0, mg_u=(50, 4, 2, 1, 30),
                 mg_d=(10, 2, 1, 1, 30), tol_u=1e-16, tol_d_rel=1e-16, tol_d_abs=1e-16)
    o.image_mp(img, [0, 1], [3e9, 70e9], [0.3, 0.2], [100.0, 1000.0], 1e-6)
    for f in (O.K, O.G, O.GC):
        o.coarse_vr_nodal(f)
    ref = staggered(o, steps, 1e-6, 2.0, schedule=schedule, du=2.0 if schedule == "ue-multiple" else 0.05)
    for g, e in zip(got, ref):
        say(f"  T={g['T']:.0f} U {g['U']:.6e}/{e['Ut']:.6e} Fw {g['Fw']:.6e}/{e['Fw']:.6e}")
        assert g["U"] == pytest.approx(e["Ut"], rel=1e-9)
        assert g["Fw"] == pytest.approx(e["Fw"], rel=1e-7)
        assert g["delatad_max"] == pytest.approx(e["ddmax"], rel=1e-7, abs=1e-12)


def test_driver_resume_reproduces_uninterrupted_run(tmp_path):
    """History of step 1 (u, v, w, H, d in the reference's write_vector_3D
    layout, plus the loop state), then a second process resumes from it: the
    F_U lines of steps 2 and 3 are identical to those of the run that never
    stopped (the device path is deterministic)."""
    require_gpu()
    from hpcrack_amd import io as hio
    write_case(tmp_path)
    hist = tmp_path / "hist"
    say("[gpu] driver history + resume")
    # fixed load increments of 2 Ue: H and d grow from step 0
    r = subprocess.run(driver_args(tmp_path, 4, "full.dat") + ["--du-ue", "2", "--write-history",
                                                                "--history-every", "1", "--history-dir", str(hist)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for t in range(4):
        for name in ("u", "v", "w", "H", "d"):
            assert os.path.getsize(hist / f"{name}_t_{t}_Proc_0.dat") == 17 ** 3 * 8
    d1 = hio.read_vector_3d(str(hist), "d", 1, 0, (17, 17, 17))
    H1 = hio.read_vector_3d(str(hist), "H", 1, 0, (17, 17, 17))
    assert H1.max() > 0 and np.isfinite(d1).all()
    st = hio.read_step_state(str(hist), 1)
    assert st["t"] == 1 and st["f_norm"] > 0
    full = open(tmp_path / "full.dat").read().splitlines()
    (tmp_path / "resumed.dat").write_text("\n".join(full[:2]) + "\n")
    r = subprocess.run(driver_args(tmp_path, 4, "resumed.dat") + ["--du-ue", "2", "--resume", "1",
                                                                   "--history-dir", str(hist)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    resumed = open(tmp_path / "resumed.dat").read().splitlines()
    assert resumed == full


def test_driver_history_files_match_oracle_fields(tmp_path):
    """The history the driver writes after each step (u, v, w, H, d in the
    write_vector_3D layout, write_read_history.cpp:18-53) against the
    oracle's fields after the same staggered steps (2 Ue increments: H and d
    grow from step 0; steps 0-2, before the specimen is nearly cut)."""
    require_gpu()
    from hpcrack_amd import io as hio
    img = write_case(tmp_path)
    hist = tmp_path / "hist"
    steps = 3
    say("[gpu] driver history files vs oracle fields")
    r = subprocess.run(driver_args(tmp_path, steps) + ["--du-ue", "2", "--write-history", "--history-every", "1",
                                                        "--history-dir", str(hist)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    o = O.Oracle(2, (4, 4, 4), (0, 16, 0, 16, 0, 16), small_k=1e-6, lc=2.0, mg_u=(50, 4, 2, 1, 30),
                 mg_d=(10, 2, 1, 1, 30), tol_u=1e-16, tol_d_rel=1e-16, tol_d_abs=1e-16)
    o.image_mp(img, [0, 1], [3e9, 70e9], [0.3, 0.2], [100.0, 1000.0], 1e-6)
    for f in (O.K, O.G, O.GC):
        o.coarse_vr_nodal(f)
    ref = staggered(o, steps, 1e-6, 2.0, schedule="ue-multiple", du=2.0, snapshots=True)
    dims = (17, 17, 17)
    for t, e in enumerate(ref):
        u = np.stack([hio.read_vector_3d(str(hist), n, t, 0, dims) for n in ("u", "v", "w")])
        H = hio.read_vector_3d(str(hist), "H", t, 0, dims)
        d = hio.read_vector_3d(str(hist), "d", t, 0, dims)
        eu = rel_l2_mod_rotation(u, e["u"], dims, (0, 16, 0, 16, 0, 16))
        eh, ed = rel_l2(H, e["H"]), rel_l2(d, e["d"])
        say(f"  t={t}: u (modulo rotation) {eu:.2e}, H {eh:.2e}, d {ed:.2e}, d_max {d.max():.4f}")
        assert eu < 1e-9 and eh < 1e-9 and ed < 1e-9
    assert ref[-1]["d"].max() > 0.5
