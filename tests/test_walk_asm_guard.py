"""Build guard for the hand-written Barnes-Hut walk (bh_traverse in csrc/bh.hip).

The walk's inline asm uses s20-s71 directly, including s32, and writes m0 as
the lane select of v_writelane_b32 (an SGPR data operand takes the one scalar
read a VOP3 instruction has on this ISA). The compiler treats s32 (its stack
pointer) and m0 as reserved and ignores such clobbers, so the asm is correct
only while:

* the kernel has no private segment: ScratchSize 0, no dynamic stack, no calls
  (s32 is then never a stack pointer), and
* no compiler-generated instruction of the kernel reads or writes m0 or s32
  (nothing the compiler keeps there can be overwritten by the asm).

This test compiles csrc/bh.hip for gfx950 with the library's flags and checks
both in the resource remarks and the AMD assembly of that build. Run on the
CPU; it fails on any change that breaks either condition.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nbody-gpu-sim_amd")
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", f"-I{ROOT}/include",
         f"-I{PKG}/csrc", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
KERNEL = "bh_traverse"


@pytest.fixture(scope="module")
def build(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("walk")
    asm = out / "bh.s"
    p = subprocess.run([HIPCC, *FLAGS, "--cuda-device-only", "-S", "-o", str(asm),
                        os.path.join(PKG, "csrc", "bh.hip"),
                        "-Rpass-analysis=kernel-resource-usage"],
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-3000:]
    return p.stderr, asm.read_text()


def remarks_of(stderr, kernel):
    lines = stderr.splitlines()
    for i, ln in enumerate(lines):
        if "Function Name:" in ln and kernel in ln:
            block = {}
            for ln2 in lines[i + 1:i + 40]:
                if "Function Name:" in ln2:
                    break
                m = re.search(r"remark:\s+([A-Za-z ]+?)(?: \[[^\]]*\])?:\s*(\S+)", ln2)
                if m:
                    block[m.group(1).strip()] = m.group(2)
            return block
    raise AssertionError(f"no resource remarks for {kernel}")


def kernel_body(asm, kernel):
    start = None
    lines = asm.splitlines()
    for i, ln in enumerate(lines):
        if re.match(r"^_Z\S*" + kernel + r"\S*:\s*(;.*)?$", ln):
            start = i
            break
    assert start is not None, f"{kernel} not in the assembly"
    body = []
    for ln in lines[start + 1:]:
        if ln.strip().startswith(".Lfunc_end"):
            break
        body.append(ln)
    return body


def test_walk_has_no_private_segment(build):
    stderr, _ = build
    r = remarks_of(stderr, KERNEL)
    assert r.get("ScratchSize") == "0", r
    assert r.get("Dynamic Stack") == "False", r
    assert r.get("SGPRs Spill") == "0" and r.get("VGPRs Spill") == "0", r
    # the walk is tuned for 8 resident waves per SIMD (at most 80 SGPRs)
    assert int(r["TotalSGPRs"]) <= 80, r


def test_no_compiler_code_uses_m0_or_s32(build):
    _, asm = build
    body = kernel_body(asm, KERNEL)
    inside, outside = False, []
    n_asm = 0
    for ln in body:
        s = ln.strip()
        if s.startswith(";;#ASMSTART"):
            inside, n_asm = True, n_asm + 1
            continue
        if s.startswith(";;#ASMEND"):
            inside = False
            continue
        if not inside:
            outside.append(s.split(";")[0])
    assert n_asm >= 2, "the walk's inline asm blocks were not found"
    code = [c for c in outside if c and not c.startswith(".")]
    assert not any(re.search(r"\bs_(swappc|setpc|call)", c) for c in code), "calls in the walk"
    bad = [c for c in code if re.search(r"\bm0\b", c) or touches_s32(c)]
    assert not bad, f"compiler-generated code touches m0/s32: {bad[:5]}"


def touches_s32(instr):
    if re.search(r"\bs32\b", instr):
        return True
    return any(int(a) <= 32 <= int(b) for a, b in re.findall(r"\bs\[(\d+):(\d+)\]", instr))


def test_register_range_parser():
    assert touches_s32("s_load_dwordx4 s[32:35], s[0:1], 0x0")
    assert touches_s32("s_mov_b32 s32, 0")
    assert touches_s32("s_load_dwordx8 s[28:35], s[4:5], 0x0")
    assert not touches_s32("s_mov_b64 s[30:31], s[2:3]")
    assert not touches_s32("s_add_i32 s33, s3, 1")
