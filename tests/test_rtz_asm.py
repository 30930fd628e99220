"""The exact fixed-point sums of the mixed force kernel and the mixed adjoint
add FP32 terms to FP64 accumulators with the FP64 rounding mode set to round
toward zero for the pair loop.  Only the adds written for that (inline asm)
may run under the switched mode: an FP64 instruction emitted by the compiler
between the two mode switches would round differently from what its source
says.  This test compiles the two sources for gfx950 (device assembly only,
no GPU needed) and checks every instantiation of the kernels that switch the
mode: every switch to round-toward-zero is followed by a switch back before
the next one and the end, and between them no compiler-emitted FP64 arithmetic (conversions from FP32,
which are exact in any mode, are allowed)."""
import os
import re
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "md.error_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
KERNELS = re.compile(r"^(_ZN3sre\d+k_(force_mixed|force_cells|adjoint_pairs|adjoint_piped)\S*):")


def device_asm(src, out):
    cmd = [HIPCC, "--offload-arch=gfx950", "--cuda-device-only", "-S", "-O3", "-std=c++17",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return open(out).read().split("\n")


def check(lines):
    """Returns (kernels checked, asm FP64 ops under RTZ, problems)."""
    problems, checked, n_asm = [], 0, 0
    i = 0
    while i < len(lines):
        m = KERNELS.match(lines[i])
        if not m:
            i += 1
            continue
        name = m.group(1)
        j = i + 1
        while not lines[j].startswith(".Lfunc_end"):
            j += 1
        body = lines[i:j]
        n_set = sum("hwreg(HW_REG_MODE, 2, 2), 3" in l for l in body)
        fp64_form = "Lb1E" in name and "adjoint_pairs" in name
        if n_set == 0 and fp64_form:
            i = j
            continue
        checked += 1
        if n_set < 1:
            problems.append(f"{name}: no switch to round toward zero")
        on = in_asm = False
        for l in body:
            t = l.strip()
            if t.startswith(";;#ASMSTART"):
                in_asm = True
            elif t.startswith(";;#ASMEND"):
                in_asm = False
            elif "hwreg(HW_REG_MODE, 2, 2), 3" in t:
                if on:
                    problems.append(f"{name}: switched to round toward zero twice")
                on = True
            elif "hwreg(HW_REG_MODE, 2, 2), 0" in t:
                on = False
            elif on and re.match(r"v_\w*f64", t):
                if in_asm:
                    n_asm += 1
                elif not t.startswith("v_cvt_f64_f32"):
                    problems.append(f"{name}: compiler FP64 op under round toward zero: {t}")
        if on:
            problems.append(f"{name}: round toward zero not switched back")
        i = j
    return checked, n_asm, problems


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
@pytest.mark.parametrize("src,min_kernels", [("k_force.hip", 100), ("k_adjoint.hip", 20)])
def test_round_toward_zero_region_holds_only_the_exact_adds(src, min_kernels):
    with tempfile.TemporaryDirectory() as d:
        lines = device_asm(src, os.path.join(d, "out.s"))
    checked, n_asm, problems = check(lines)
    assert not problems, problems[:10]
    assert checked >= min_kernels and n_asm > 0
