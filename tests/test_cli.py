"""The command-line front end (cuszp_amd/cuszp, options of the reference's
examples/cuSZp.cpp): argument handling on the CPU; on the GPU, the stream it
writes (-x) and the values it writes (-o) against the oracle, absolute and
relative bounds."""
import os
import subprocess

import numpy as np
import pytest

from tests.workloads import grf_volume, mixed_field

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cuszp_amd", "cuszp")


def _run(args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)


def test_cli_usage_and_argument_errors():
    if not os.path.exists(CLI):
        pytest.skip("cuszp_amd/cuszp not built")
    r = _run([])
    assert r.returncode == 1 and "Usage:" in r.stdout
    r = _run(["-i", "x", "-t", "f16"])
    assert r.returncode == 1 and "Unsupported data type" in r.stdout
    r = _run(["-i", "x", "-t", "f32", "-m", "plain", "-q"])
    assert r.returncode == 1 and "Unrecognized option" in r.stdout
    r = _run(["-i", "x", "-t", "f32", "-m", "plain", "-d", "1"])
    assert r.returncode == 1 and "Missing required arguments" in r.stdout
    r = _run(["-i", "x", "-t", "f32", "-m", "plain", "-d", "4", "-eb", "abs", "1e-3"])
    assert r.returncode == 1 and "Unsupported processing dimension" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    ("f32", "plain", 1, None, "abs", 1e-3),
    ("f32", "outlier", 3, (64, 48, 40), "rel", 1e-4),
    ("f64", "fixed", 2, (96, 30, 3), "abs", 1e-2),
    ("f64", "outlier", 1, None, "rel", 1e-3),
])
def test_cli_matches_oracle(device, oracle, tmp_path, case):
    typ, mode_name, dim, dims, ebmode, ebval = case
    dtype = np.float32 if typ == "f32" else np.float64
    mode = {"fixed": 0, "plain": 1, "outlier": 2}[mode_name]
    if dims is None:
        x = mixed_field(250_001, seed=3, dtype=dtype)
        dargs, odims = ["-d", "1"], (0, 0, 0)
    else:
        xd, yd, zd = dims
        x = grf_volume(zd, yd, xd, seed=4, dtype=dtype).ravel()
        dargs, odims = ["-d", str(dim), str(zd), str(yd), str(xd)], (xd, yd, zd)
    src, cmpf, decf = tmp_path / "in.raw", tmp_path / "out.cmp", tmp_path / "out.dec"
    x.tofile(src)
    r = _run(["-i", str(src), "-t", typ, "-m", mode_name, *dargs, "-eb", ebmode, str(ebval),
              "-x", str(cmpf), "-o", str(decf)])
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "Pass error check!" in r.stdout
    eb = float(np.float32(ebval)) if ebmode == "abs" else oracle.rel_eb(x, ebval)[0]
    printed = float(r.stdout.split("error bound:")[1].split()[0])
    assert np.float32(printed) == np.float32(eb)
    want = oracle.compress(x, eb, mode, dim, odims)
    assert cmpf.read_bytes() == want
    y = np.fromfile(decf, dtype=dtype)
    yref = oracle.decompress(want, x.size, dtype, eb, mode, dim, odims)
    assert np.array_equal(y.view(np.uint8), yref.view(np.uint8))
