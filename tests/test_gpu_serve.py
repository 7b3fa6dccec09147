"""The request server (hpcrack_amd/hpcrack_serve): continuous batching of
independent single-increment fracture solves under an HBM budget.  Each
request's result must be identical whether it ran alone (one slot, nothing
else on the device) or in the mix (several admitted requests overlapping on
their own streams), for the FP64 and the mixed-precision forms; the byte
estimate used for admission must match what a context holds."""
import json
import os
import subprocess

import pytest

from gpu_common import require_gpu, say

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVER = os.path.join(ROOT, "hpcrack_amd", "hpcrack_serve")


def results(path):
    out = {}
    for line in open(path):
        f = line.split()
        kv = dict(zip(f[0::2], f[1::2]))
        kv.pop("latency_ms")
        out[int(kv["id"])] = kv
    return out


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_results_identical_alone_and_in_the_mix(tmp_path, mixed):
    require_gpu()
    base = [SERVER, "--requests", "10", "--rate", "400", "--sizes", "17,33", "--seed", "7", "--budget-gb", "50"]
    base.append("--mixed" if mixed else "--fp64")
    runs = {}
    for slots in (4, 1):
        out = tmp_path / f"res_{slots}.txt"
        say(f"[gpu] server {'mixed' if mixed else 'fp64'} slots={slots}")
        r = subprocess.run(base + ["--slots", str(slots), "--out", str(out)], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        summary = json.loads(r.stdout.strip().splitlines()[-1])
        say(f"  {summary}")
        assert summary["failed"] == 0 and summary["requests"] == 10
        assert summary["latency_ms_p99"] >= summary["latency_ms_p50"] > 0
        runs[slots] = (summary, results(out))
    assert runs[4][0]["peak_concurrent"] > 1
    assert runs[1][0]["peak_concurrent"] == 1
    mix, alone = runs[4][1], runs[1][1]
    assert mix.keys() == alone.keys()
    for i in mix:
        assert mix[i] == alone[i], (i, mix[i], alone[i])
        assert int(mix[i]["rc"]) == 0
        if mixed:
            # the flexible CG reaches the tolerance; the FAS cycle may stop at
            # its cycle cap above it (the reference's behaviour)
            assert float(mix[i]["rel_u"]) <= 1e-16


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_fused_serving_identical_to_requests_alone(tmp_path, mixed):
    """--fuse: the cycles of concurrently admitted requests run as
    multi-problem launches through one fusion engine (the mixed form fuses its
    phase-field cycles); every request's result is that of the request alone."""
    require_gpu()
    base = [SERVER, "--requests", "12", "--rate", "400", "--sizes", "17,33", "--seed", "11", "--budget-gb", "50"]
    base.append("--mixed" if mixed else "--fp64")
    runs = {}
    for slots, extra in ((6, ["--fuse", "2000"]), (1, [])):
        out = tmp_path / f"res_{slots}.txt"
        say(f"[gpu] server {'mixed' if mixed else 'fp64'} slots={slots} {' '.join(extra)}")
        r = subprocess.run(base + ["--slots", str(slots), "--out", str(out)] + extra, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        summary = json.loads(r.stdout.strip().splitlines()[-1])
        say(f"  {summary}")
        assert summary["failed"] == 0 and summary["requests"] == 12
        runs[slots] = (summary, results(out))
    fused = runs[6][0]
    assert fused["fused"] is True and fused["fused_cycles"] > 0
    assert fused["fused_problem_launches"] >= fused["fused_launches"] > 0
    mix, alone = runs[6][1], runs[1][1]
    assert mix.keys() == alone.keys()
    for i in mix:
        assert mix[i] == alone[i], (i, mix[i], alone[i])


def test_byte_estimate_matches_context():
    HA = require_gpu()
    # the last one has a finest level of split-output pass A tiles (side arrays)
    for ml, enb in ((2, (4, 4, 4)), (3, (4, 8, 4)), (1, (2, 6, 4)), (1, (64, 64, 4))):
        dom = (0.0, enb[0] * 2.0 ** ml, 0.0, enb[1] * 2.0 ** ml, 0.0, enb[2] * 2.0 ** ml)
        s = HA.Solver(ml, enb, dom, tol_u=1e-16)
        assert s.device_bytes() == HA.estimate_bytes(ml, enb, dom)
        for f, v in ((HA.K, 2.5e3), (HA.G, 1.2e3), (HA.GC, 100.0)):
            for l in range(ml + 1):
                s.fill(f, l, v, 0)
        s.u_entire_mixed(1.0, 0, 0.0, max_iter=2)
        assert s.device_bytes() == HA.estimate_bytes(ml, enb, dom, mixed=True)
