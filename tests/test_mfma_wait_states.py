"""Wait states around the matrix-core products of the harmonic kernel.

The planar order sums of the matrix-core variants are asm statements of
v_mfma_f32_4x4x1_16b_f32 with the accumulators tied in AGPRs; the compiler
does not see into them, so it cannot keep the wait an MFMA result needs
before other instructions touch its registers.  This compiles each order
bucket's matrix-core object (the Makefile's flags) to assembly and checks
that no instruction reads or writes an AGPR that a statement's MFMAs wrote
within 8 wait states after the statement (the chunk sizes whose sums spill
did exactly that and were withdrawn).  The one exception is the next
statement's own products, which add to the same accumulators (destination
and addend the same registers): they keep at least the 2 wait states that a
statement keeps between its cos and sin products on one accumulator.  CPU
only: hipcc cross-compiles.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bimsim-cuda_amd")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
BUCKETS = (16, 24, 32, 40, 48, 64, 80, 96, 112)
WAIT = 8
WAIT_ACCUMULATE = 2

FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Xclang", "-target-feature",
         "-Xclang", "-packed-fp32-ops", "-ffp-contract=off", "-mllvm", "-disable-machine-licm",
         "-DHARM_PART=2", "-S", "--cuda-device-only"]


def agprs(text):
    """AGPR numbers named in an instruction (a7, a[4:7])."""
    out = set()
    for lo, hi in re.findall(r"\ba\[(\d+):(\d+)\]", text):
        out.update(range(int(lo), int(hi) + 1))
    for r in re.findall(r"\ba(\d+)\b", text):
        out.add(int(r))
    return out


def early_uses(asm):
    """(kernel, wait states, instruction) for every access to an AGPR that an
    MFMA asm statement wrote, within WAIT wait states after the statement."""
    hits, statements = [], 0
    for blk in re.split(r"\n(?=_ZN6bimsim4harm18series_harm_kernel)", "\n" + asm)[1:]:
        name = blk.split(":", 1)[0]
        lines = [l.strip() for l in blk.split("\n")]
        for i, l in enumerate(lines):
            if l != ";;#ASMEND":
                continue
            j, written = i - 1, set()
            while j > 0 and lines[j] != ";;#ASMSTART":
                m = re.match(r"v_mfma\S*\s+(a\[\d+:\d+\])", lines[j])
                if m:
                    written |= agprs(m.group(1))
                j -= 1
            if not written:
                continue
            statements += 1
            ws = 0
            for x in lines[i + 1:]:
                if ws >= WAIT:
                    break
                if not x or x.startswith(";") or x.startswith("."):
                    continue
                if x.startswith("s_nop"):
                    ws += int(x.split()[1]) + 1
                    continue
                if agprs(x) & written:
                    acc = re.match(r"v_mfma\S*\s+(a\[\d+:\d+\]),\s*v\d+,\s*v\d+,\s*(a\[\d+:\d+\])", x)
                    if not (acc and acc.group(1) == acc.group(2) and ws >= WAIT_ACCUMULATE):
                        hits.append((name[:80], ws, x))
                ws += 1
    return statements, hits


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
@pytest.mark.parametrize("L", BUCKETS)
def test_no_accumulator_access_within_the_mfma_wait(L, tmp_path):
    out = tmp_path / f"harm_mfma_{L}.s"
    r = subprocess.run([HIPCC, *FLAGS, f"-DHARM_L={L}", "csrc/us_harm.hip", "-o", str(out)],
                       cwd=PKG, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    statements, hits = early_uses(out.read_text())
    assert statements > 0, "no MFMA asm statements found"
    assert hits == [], hits[:5]


def test_the_check_finds_an_early_read():
    """The scan itself: a read of a written accumulator right after the
    statement is found; one after `s_nop 7` is not."""
    asm = "\n".join([
        "_ZN6bimsim4harm18series_harm_kernelILi40E: ; kernel",
        ";;#ASMSTART", "s_nop 1",
        "v_mfma_f32_4x4x1_16b_f32 a[16:19], v4, v29, a[16:19] cbsz:4 abid:6",
        "s_nop 0", ";;#ASMEND",
        "v_accvgpr_read_b32 v27, a19",
        ";;#ASMSTART", "s_nop 1",
        "v_mfma_f32_4x4x1_16b_f32 a[0:3], v4, v29, a[0:3] cbsz:4",
        "s_nop 0", ";;#ASMEND",
        "s_nop 7",
        "v_accvgpr_read_b32 v27, a3",
    ])
    statements, hits = early_uses(asm)
    assert statements == 2
    assert [(w, x) for _, w, x in hits] == [(0, "v_accvgpr_read_b32 v27, a19")]


def test_the_check_allows_the_next_products_on_the_sums():
    """The next statement's products on the same accumulators pass after 2
    wait states (counted from the statement's end) and are found with none;
    a product into other registers that reads the sums as its addend is
    found."""
    def stmt(*mfma):
        return [";;#ASMSTART", "s_nop 1", *mfma, "s_nop 0", ";;#ASMEND"]
    head = ["_ZN6bimsim4harm18series_harm_kernelILi16E: ; kernel"]
    first = stmt("v_mfma_f32_4x4x1_16b_f32 a[0:3], v2, v3, a[0:3] cbsz:4 abid:2")
    ok = head + first + stmt("v_mfma_f32_4x4x1_16b_f32 a[0:3], v10, v4, a[0:3] cbsz:4 abid:2")
    assert early_uses("\n".join(ok))[1] == []
    early = head + first + [";;#ASMSTART", "v_mfma_f32_4x4x1_16b_f32 a[0:3], v10, v4, a[0:3]",
                            ";;#ASMEND"]
    assert [w for _, w, _ in early_uses("\n".join(early))[1]] == [0]
    other = head + first + stmt("v_mfma_f32_4x4x1_16b_f32 a[4:7], v10, v4, a[0:3] cbsz:4 abid:2")
    assert [w for _, w, _ in early_uses("\n".join(other))[1]] == [2]
