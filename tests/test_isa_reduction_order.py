"""The reduction ticket's memory ordering in the built gfx950 code.

Each block of a reducing kernel stores its partial sums with agent-scope
atomic stores (`global_store_* ... sc1`, written through to the L2), waits
for their completion, and only then takes its completion ticket
(`global_atomic_add`); the last block reads the partials after the last
ticket.  Without the release fence (the default) that order is not a
memory-model guarantee of the source: it rests on the compiler emitting
store -> s_waitcnt vmcnt(0) -> atomic.  This test compiles the kernels to
AMD assembly and checks, on the control-flow graph of every kernel that
takes a ticket, that every path from a partial store to the ticket passes a
vmcnt(0) wait.  CPU only (hipcc cross-compiles)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["hpcr_kernels.hip", "hpcr_mixed.hip"]


def kernels(asm):
    """{symbol: [instruction lines]} of the kernel bodies"""
    out, cur = {}, None
    for line in asm.splitlines():
        m = re.match(r"^(_Z\S+):", line)
        if m:
            cur = m.group(1)
            out[cur] = []
            continue
        if cur is None:
            continue
        t = line.split(";")[0].strip()
        if t.startswith(".Lfunc_end"):
            cur = None
            continue
        if t:
            out[cur].append(t)
    return out


def blocks(body):
    """basic blocks: (label or None, [instructions]) in text order"""
    bl, cur = [], [None, []]
    for t in body:
        if re.match(r"^\.LBB\w+:$", t):
            bl.append(cur)
            cur = [t[:-1], []]
        elif t.startswith("."):
            continue
        else:
            cur[1].append(t)
            if t.startswith(("s_branch", "s_cbranch", "s_endpgm")):
                bl.append(cur)
                cur = [None, []]
    bl.append(cur)
    return [b for b in bl if b[0] is not None or b[1]]


def successors(bl, i):
    label = {b[0]: k for k, b in enumerate(bl) if b[0]}
    ins = bl[i][1]
    last = ins[-1] if ins else ""
    succ = []
    if last.startswith(("s_branch", "s_cbranch")):
        succ.append(label[last.split()[1]])
    if not last.startswith(("s_branch", "s_endpgm")) and i + 1 < len(bl):
        succ.append(i + 1)
    return succ


def is_wait0(t):
    return t.startswith("s_waitcnt") and ("vmcnt(0)" in t or t == "s_waitcnt 0")


def unordered_paths(body):
    """number of (partial store, ticket) pairs joined by a path without a
    vmcnt(0) wait; and the number of tickets checked"""
    bl = blocks(body)
    bad, tickets = 0, 0
    for i, (_, ins) in enumerate(bl):
        for j, t in enumerate(ins):
            if not (t.startswith("global_store") and t.endswith("sc1")):
                continue
            # walk forward from the store; a path is cut at a vmcnt(0) wait
            seen = set()
            stack = [(i, j + 1)]
            while stack:
                b, k = stack.pop()
                if (b, k) in seen:
                    continue
                seen.add((b, k))
                cut = False
                for t2 in bl[b][1][k:]:
                    if is_wait0(t2):
                        cut = True
                        break
                    if t2.startswith("global_atomic_add"):
                        bad += 1
                        cut = True
                        break
                if not cut:
                    stack.extend((s, 0) for s in successors(bl, b))
    for _, ins in bl:
        tickets += sum(1 for t in ins if t.startswith("global_atomic_add"))
    return bad, tickets


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("isa")
    out = {}
    for src in SOURCES:
        s = os.path.join(d, src + ".s")
        cmd = [HIPCC, "-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", "--cuda-device-only", "-S", "-o", s, "-I", os.path.join(ROOT, "include"),
               os.path.join(ROOT, "hpcrack_amd", "csrc", src)]
        subprocess.run(cmd, check=True, capture_output=True)
        out[src] = open(s).read()
    return out


def test_partials_complete_before_every_ticket(asm):
    checked = 0
    for src, text in asm.items():
        for name, body in kernels(text).items():
            if not any(t.startswith("global_atomic_add") for t in body):
                continue
            if not any(t.startswith("global_store") and t.endswith("sc1") for t in body):
                continue
            bad, tickets = unordered_paths(body)
            assert bad == 0, f"{src}: {name}: a partial store reaches the ticket without a vmcnt(0) wait"
            checked += tickets
    # the FP64 passes, the preconditioner passes and the pointwise reductions
    assert checked >= 20


def test_checker_finds_a_missing_wait():
    """The path check itself: a store that reaches the ticket unwaited on one
    branch is reported, the waited form is not."""
    good = ["global_store_dwordx2 v4, v[2:3], s[2:3] sc1", "s_cbranch_scc1 .LBB0_2", "buffer_wbl2 sc1",
            "s_waitcnt vmcnt(0)", "s_branch .LBB0_3", ".LBB0_2:", "s_waitcnt vmcnt(0) expcnt(0) lgkmcnt(0)",
            ".LBB0_3:", "global_atomic_add v3, v3, v4, s[16:17] sc0", "s_endpgm"]
    bad = ["global_store_dwordx2 v4, v[2:3], s[2:3] sc1", "s_cbranch_scc1 .LBB0_2", "buffer_wbl2 sc1",
           "s_waitcnt vmcnt(0)", "s_branch .LBB0_3", ".LBB0_2:", "s_waitcnt lgkmcnt(0)",
           ".LBB0_3:", "global_atomic_add v3, v3, v4, s[16:17] sc0", "s_endpgm"]
    assert unordered_paths(good) == (0, 1)
    assert unordered_paths(bad)[0] == 1
