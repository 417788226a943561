"""ISA-level guard for the hand-written memory instructions (no GPU needed).

The hot loops issue some stores through inline asm. A gfx9 VMEM instruction
that reads an SGPR written by a VALU instruction (for instance the
v_readfirstlane that makes a base address wave-uniform) needs 5 wait states,
and the compiler's hazard recognizer does not treat an inline asm block as a
VMEM instruction. This test compiles every kernel translation unit to gfx950
assembly and checks that no store with an SGPR base follows a VALU write of
that SGPR within 5 wait states.
"""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "cuszp_amd"))

import build as B  # noqa: E402

UNITS = ["kern_f32.hip", "kern_f64.hip", "kern_batch_c.hip", "kern_batch_d.hip"]


def _asm(unit, tmp_path):
    out = tmp_path / (unit + ".s")
    cmd = [B.HIPCC, *B.CFLAGS, "-x", "hip", "-S", "--offload-device-only", os.path.join(B.CSRC, unit),
           "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return out.read_text()


def sgpr_vmem_hazards(asm: str):
    """Stores reading an SGPR base written by a VALU op fewer than 5 wait states earlier."""
    ins = []
    for line in asm.split("\n"):
        t = line.strip()
        if not t or t.startswith((";", ".")) or t.endswith(":"):
            continue
        ins.append(t)
    found = []
    for i, t in enumerate(ins):
        m = re.match(r"(global|buffer)_store_\w+ .*s\[(\d+):(\d+)\]", t)
        if not m:
            continue
        regs = set(range(int(m.group(2)), int(m.group(3)) + 1))
        waits = 0
        for j in range(1, 12):
            p = ins[i - j]
            one = re.match(r"v_\w+\s+s(\d+)\b", p)
            two = re.match(r"v_\w+\s+s\[(\d+):(\d+)\]", p)
            written = set()
            if one:
                written = {int(one.group(1))}
            elif two:
                written = set(range(int(two.group(1)), int(two.group(2)) + 1))
            if written & regs:
                found.append((waits, p, t))
                break
            n = re.match(r"s_nop (\d+)", p)
            waits += int(n.group(1)) + 1 if n else 1
            if waits >= 5:
                break
    return found


def test_scanner_flags_the_hazard():
    asm = "\tv_readfirstlane_b32 s6, v2\n\tv_readfirstlane_b32 s7, v3\n\tglobal_store_dword v1, v4, s[6:7]\n"
    assert len(sgpr_vmem_hazards(asm)) == 1
    safe = asm.replace("\tglobal_store", "\ts_nop 4\n\tglobal_store")
    assert sgpr_vmem_hazards(safe) == []


@pytest.mark.slow
def test_no_valu_sgpr_to_vmem_hazard(tmp_path):
    if not os.path.exists(B.HIPCC):
        pytest.skip("hipcc not available")
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        asms = list(ex.map(lambda u: _asm(u, tmp_path), UNITS))
    for unit, asm in zip(UNITS, asms):
        bad = sgpr_vmem_hazards(asm)
        assert not bad, (unit, bad[:5])
