"""Self-describing container (include/cuSZp.h): a 64-byte header in front of
the unchanged raw stream.

CPU: the header the library writes is checked field by field with Python's
struct module and zlib's CRC-32 (an implementation independent of the
library's), and every corruption the reader must refuse is refused.
GPU: cuSZp_compress_container / cuSZp_decompress_container round trips whose
raw stream equals the oracle's and whose values equal the oracle's
decompression; the CLI's -X / -r options.
"""
import ctypes
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

from tests.workloads import grf_volume, mixed_field

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cuszp_amd", "cuszp")
MAGIC = b"\x89CSZP\r\n\x1a"
CORRUPT, INVALID_ARG, INVALID_DIMS = 5, 1, 2


@pytest.fixture(scope="module")
def L():
    from cuszp_amd import _lib

    return _lib.load()


def _info(**kw):
    from cuszp_amd._lib import ContainerInfo, Uint3

    ci = ContainerInfo()
    ci.dim, ci.type, ci.mode = kw.get("dim", 1), kw.get("type", 0), kw.get("mode", 1)
    ci.relative = kw.get("relative", 0)
    ci.dims = Uint3(*kw.get("dims", (0, 0, 0)))
    ci.nbEle = kw.get("n", 100_000)
    ci.errorBound = kw.get("eb", 1e-3)
    ci.relBound = kw.get("rel", 0.0)
    ci.streamBytes = kw.get("stream", 5000)
    return ci


def _write(L, ci):
    buf = (ctypes.c_ubyte * 64)()
    rc = L.cuSZp_container_write_header(ctypes.byref(ci), buf)
    return rc, bytes(buf)


def _read(L, data):
    from cuszp_amd._lib import ContainerInfo

    ci = ContainerInfo()
    buf = (ctypes.c_ubyte * max(len(data), 1)).from_buffer_copy(data or b"\0")
    return L.cuSZp_container_read_header(buf, len(data), ctypes.byref(ci)), ci


def _expected_header(dim, typ, mode, rel, dims, n, eb, relb, stream):
    body = struct.pack("<8sHHBBBBIIIIQddQ", MAGIC, 1, 64, dim, typ, mode, rel, *dims, 0, n, eb, relb, stream)
    crc = zlib.crc32(body)
    return body[:28] + struct.pack("<I", crc) + body[32:]


@pytest.mark.parametrize("kw", [
    dict(dim=1, type=0, mode=1, n=100_000, eb=1e-3, stream=5000),
    dict(dim=3, type=1, mode=2, dims=(64, 48, 40), n=64 * 48 * 40, eb=2.5e-4, relative=1, rel=1e-4, stream=9000),
    dict(dim=2, type=0, mode=0, dims=(96, 30, 3), n=96 * 30 * 3, eb=0.01, stream=200),
])
def test_header_layout(L, kw):
    ci = _info(**kw)
    rc, h = _write(L, ci)
    assert rc == 0
    eb = float(np.float32(kw["eb"])) if kw.get("type", 0) == 0 else kw["eb"]
    want = _expected_header(kw["dim"], kw["type"], kw["mode"], kw.get("relative", 0), kw.get("dims", (0, 0, 0)),
                            kw["n"], eb, kw.get("rel", 0.0), kw["stream"])
    assert h == want
    rc, back = _read(L, h + b"\0" * 10)
    assert rc == 0
    assert (back.dim, back.type, back.mode, back.relative) == (kw["dim"], kw["type"], kw["mode"], kw.get("relative", 0))
    assert (back.dims.x, back.dims.y, back.dims.z) == kw.get("dims", (0, 0, 0))
    assert (back.nbEle, back.errorBound, back.relBound, back.streamBytes) == (kw["n"], eb, kw.get("rel", 0.0),
                                                                              kw["stream"])


def test_header_rejects(L):
    rc, h = _write(L, _info())
    assert rc == 0
    assert _read(L, h[:63])[0] == CORRUPT                      # short
    for pos in range(64):                                      # any flipped bit: magic, fields or CRC
        bad = bytearray(h)
        bad[pos] ^= 0x10
        assert _read(L, bytes(bad))[0] == CORRUPT, pos
    # a well-formed header (CRC correct) with a field the format forbids
    def patched(off, fmt, val):
        b = bytearray(h)
        struct.pack_into(fmt, b, off, val)
        struct.pack_into("<I", b, 28, 0)
        struct.pack_into("<I", b, 28, zlib.crc32(bytes(b)))
        return bytes(b)
    assert _read(L, patched(8, "<H", 2))[0] == CORRUPT          # version 2
    assert _read(L, patched(12, "<B", 4))[0] == CORRUPT         # dim 4
    assert _read(L, patched(14, "<B", 3))[0] == CORRUPT         # mode 3
    assert _read(L, patched(15, "<B", 2))[0] == CORRUPT         # unknown flag
    assert _read(L, patched(40, "<d", -1.0))[0] == CORRUPT      # negative bound
    assert _read(L, patched(56, "<Q", 3000))[0] == CORRUPT      # shorter than the 1D header (4096 bytes)
    assert _read(L, patched(56, "<Q", 1 << 40))[0] == CORRUPT   # longer than any stream of 100,000 values
    assert _read(L, patched(16, "<I", 8))[0] == CORRUPT         # 1D with dims
    # the writer refuses what the reader would
    assert _write(L, _info(mode=5))[0] == INVALID_ARG
    assert _write(L, _info(eb=0.0))[0] == INVALID_ARG
    assert _write(L, _info(relative=1, rel=0.0))[0] == INVALID_ARG
    assert _write(L, _info(dim=2, dims=(6, 4, 1), n=24))[0] == INVALID_DIMS   # x % 4 != 0
    assert _write(L, _info(dim=2, dims=(8, 4, 1), n=33))[0] == INVALID_DIMS   # nbEle != x*y*z


def test_cli_container_usage():
    if not os.path.exists(CLI):
        pytest.skip("cuszp_amd/cuszp not built")
    r = subprocess.run([CLI, "-h"], capture_output=True, text=True, timeout=60)
    assert "-X" in r.stdout and "-r" in r.stdout


CASES = [
    (1, None, np.float32, 1, 1e-3, 0.0),
    (1, None, np.float64, 2, 1e-4, 0.0),
    (2, (1028, 300, 2), np.float32, 0, 1e-2, 0.0),
    (3, (132, 130, 67), np.float64, 1, 1e-3, 0.0),
    (3, (64, 48, 40), np.float32, 2, None, 1e-4),   # relative bound recorded in the header
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}D-{np.dtype(c[2]).name}-m{c[3]}" for c in CASES])
def test_container_round_trip(device, oracle, case):
    import torch
    from cuszp_amd import pycuSZp as P

    dim, dims, dtype, mode, eb, rel = case
    if dim == 1:
        x = mixed_field(400_003, seed=11, dtype=dtype)
        dims = (0, 0, 0)
    else:
        x = grf_volume(dims[2], dims[1], dims[0], seed=12, dtype=dtype).ravel()
    xd = torch.from_numpy(x).to(device)
    if rel:
        eb = P.rel_to_abs(xd, rel)
    print(f"container {dim}D {np.dtype(dtype).name} mode {mode} eb {eb}", flush=True)
    box, size = P.compress_container(xd, eb, mode, dim, dims, rel_bound=rel)
    host = box[:size].cpu().numpy().tobytes()
    want = oracle.compress(x, eb, mode, dim, dims)
    assert host[64:] == want
    info = P.read_container_header(host)
    typ = 1 if dtype == np.float64 else 0
    eb_stored = float(np.float32(eb)) if typ == 0 else eb
    assert info == {"dim": dim, "type": typ, "mode": mode, "relative": rel > 0, "dims": tuple(dims),
                    "num_elements": x.size, "error_bound": eb_stored, "rel_bound": rel, "stream_bytes": len(want)}
    y, info2 = P.decompress_container(box, size)
    assert info2 == info and y.dtype == xd.dtype and y.numel() == x.size
    yref = oracle.decompress(want, x.size, dtype, eb, mode, dim, dims)
    iv = np.uint64 if typ else np.uint32
    assert np.array_equal(y.cpu().numpy().view(iv), yref.view(iv))
    # a container one byte short of its stream is refused before any launch
    with pytest.raises(P.CuSZpError):
        P.decompress_container(box, size - 1)


@pytest.mark.gpu
def test_cli_container_matches_oracle(device, oracle, tmp_path):
    if not os.path.exists(CLI):
        pytest.skip("cuszp_amd/cuszp not built")
    x = grf_volume(40, 48, 64, seed=4, dtype=np.float32).ravel()
    src, box, dec = tmp_path / "in.raw", tmp_path / "out.cszp", tmp_path / "out.dec"
    x.tofile(src)
    r = subprocess.run([CLI, "-i", str(src), "-t", "f32", "-m", "outlier", "-d", "3", "40", "48", "64",
                        "-eb", "rel", "1e-4", "-X", str(box)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "Pass error check!" in r.stdout
    eb = oracle.rel_eb(x, 1e-4)[0]
    data = box.read_bytes()
    want = oracle.compress(x, eb, 2, 3, (64, 48, 40))
    assert data[64:] == want
    assert data[:64] == _expected_header(3, 0, 2, 1, (64, 48, 40), x.size, float(np.float32(eb)),
                                         float(np.float32(1e-4)), len(want))
    r = subprocess.run([CLI, "-r", str(box), "-o", str(dec), "-i", str(src)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "Pass error check!" in r.stdout and "outlier 3D" in r.stdout
    yref = oracle.decompress(want, x.size, np.float32, eb, 2, 3, (64, 48, 40))
    assert np.array_equal(np.fromfile(dec, dtype=np.float32).view(np.uint32), yref.view(np.uint32))
