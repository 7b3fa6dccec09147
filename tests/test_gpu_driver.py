"""The C++ host driver (hpcrack_amd/hpcrack_run, crack_propagation + main over
the C-ABI) against the oracle's staggered loop on a seeded 17^3 image given
as config.txt + uint8 .raw + material CSV."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from gpu_common import require_gpu, say
from test_gpu_solves import staggered

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "hpcrack_amd", "hpcrack_run")


def parse_fu(path):
    rows = []
    for line in open(path):
        f = dict(kv.split("=") for kv in line.replace(" ", "").split(","))
        rows.append({k: float(v) for k, v in f.items()})
    return rows


@pytest.mark.parametrize("schedule", ["fixed", "reference"])
def test_driver_matches_oracle_staggered_loop(tmp_path, schedule):
    require_gpu()
    from hpcrack_amd import io as hio
    n = 17
    img, _ = hio.synth_image((n, n, n), 31, n_incl=3, incl_r=(2.0, 4.0))
    hio.write_raw(str(tmp_path / "img.raw"), img)
    (tmp_path / "mat.csv").write_text("Gray level,0,1\nE/Pa,3000000000,70000000000\nnu,0.3,0.2\ngc/(N/m),100,1000\n")
    (tmp_path / "config.txt").write_text(
        "D=3\nnp=1,1,1\nX_start=0,0,0\nIMGname=%s\nvoxel_nb_X=%d\nvoxel_nb_Y=%d\nvoxel_nb_Z=%d\n"
        "ROI_start_X=0\nROI_start_Y=0\nROI_start_Z=0\nvoxel_size=1e-6\nelement_nb=4,4,4\n"
        "ncy_u=30\nnu0_u=50\nnu1_u=4\nnu2_u=2\ngamma_u=1\nncy_d=30\nnu0_d=10\nnu1_d=2\nnu2_d=1\ngamma_d=1\n"
        "lc=2\n" % (tmp_path / "img.raw", n, n, n))
    steps = 4
    args = [DRIVER, str(tmp_path / "config.txt"), "--csv", str(tmp_path / "mat.csv"), "--steps", str(steps),
            "--small-k", "1e-6", "--tol-u", "1e-16", "--tol-d", "1e-16", "--out", str(tmp_path / "F_U.dat"),
            "--stop-force", "-1e300"]
    if schedule == "fixed":
        args += ["--du", "0.05"]
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
    ref = staggered(o, steps, 1e-6, 2.0, schedule=schedule, du=0.05)
    for g, e in zip(got, ref):
        say(f"  T={g['T']:.0f} U {g['U']:.6e}/{e['Ut']:.6e} Fw {g['Fw']:.6e}/{e['Fw']:.6e}")
        assert g["U"] == pytest.approx(e["Ut"], rel=1e-9)
        assert g["Fw"] == pytest.approx(e["Fw"], rel=1e-7)
        assert g["delatad_max"] == pytest.approx(e["ddmax"], rel=1e-7, abs=1e-12)
