"""Build libcuszp_amd.so in-tree with hipcc for gfx950.

Every HIP translation unit is compiled with -ffp-contract=off (quantisation and
dequantisation must round each multiply and add separately, like the stream
format requires) and without packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcuszp_amd.so")
CLI = os.path.join(HERE, "cuszp")
SOURCES = ["kern_batch_c.hip", "kern_batch_d.hip", "kern_f32.hip", "kern_f64.hip", "kern_range.hip", "kern_quality.hip",
           "cuszp_api.cpp",
           "cuszp_batch.cpp", "cuszp_serve.cpp", "cuszp_container.cpp",
           "cuszp_noalloc.hip"]
ARCH = os.environ.get("CUSZP_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *CFLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # the command-line front end (cuszp_amd/cuszp), linked against the library
    tmp = CLI + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-x", "hip", os.path.join(HERE, "cli", "cuszp_cli.cpp"),
          "-x", "none", "-L" + HERE, "-lcuszp_amd", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
