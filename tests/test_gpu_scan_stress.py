"""Forward progress of the inter-tile scan, forced from the host.

The reference's chained scan assumes that blocks run in launch order
(src/cuSZp_kernels_1D_f32.cu:171-197: lane 0 spins on its predecessor's flag).
Here a tile that waits too long claims the pending predecessor and computes
that tile's byte count itself (compress: from the predecessor's input values;
decompress: from its header bytes), with code separate from the tile pass.
These tests make that path run on every launch:

  * hold: every 8th tile publishes nothing for 300 us (a bounded wait inside
    the kernel), far past the 100 us after which successors claim it;
  * steal at once: the claim threshold is 0, so any unresolved poll claims
    (with every 4th tile held for 20 us, so that polls find pending words).

Every stream must still be byte-identical to the oracle and every
decompression bit-identical, and the scan report must show that claims
happened (CUSZP_SCAN_INFO_STOLEN, claim count). All waits are bounded, so every
kernel exits normally.
"""
import numpy as np
import pytest

from tests.helpers import gpu_compress, gpu_decompress
from tests.workloads import grf_volume, mixed_field

pytestmark = pytest.mark.gpu

HOLD = dict(hold_every=8, hold_ticks=30_000, stats=True)   # 300 us at the 100 MHz clock
# claim at the first unresolved poll; every 4th tile also holds back for 20 us,
# so that windows stay unresolved in both directions (a decompress tile
# otherwise publishes its byte count within a microsecond of starting)
AT_ONCE = dict(steal_ticks=0, hold_every=4, hold_ticks=2_000, stats=True)

# (dim, dims, dtype, mode, eb): 20-50 tiles each, ragged edges in 2D/3D
CASES = [
    (1, None, np.float32, 1, 1e-3),
    (1, None, np.float64, 2, 1e-4),
    (2, (1028, 300, 2), np.float32, 2, 1e-2),
    (3, (132, 130, 67), np.float64, 0, 1e-3),
]


def _ids(c):
    return f"{c[0]}D-{np.dtype(c[2]).name}-m{c[3]}"


def _input(case):
    dim, dims, dtype, mode, eb = case
    if dim == 1:
        n = 600_001 if dtype == np.float32 else 300_007
        return mixed_field(n, seed=n % 89, dtype=dtype), (0, 0, 0)
    x, y, z = dims
    return grf_volume(z, y, x, seed=5, dtype=dtype).ravel(), dims


@pytest.fixture
def stress():
    """Sets scan settings for the test; restores the defaults afterwards."""
    from cuszp_amd import pycuSZp as P

    def apply(**kw):
        P.scan_config(**kw)
        P.scan_report()   # start from a clean report
    yield apply
    P.scan_config()


@pytest.mark.parametrize("knobs", [HOLD, AT_ONCE], ids=["hold", "steal-at-once"])
@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_single_array_claims(device, oracle, stress, case, knobs):
    from cuszp_amd import pycuSZp as P

    dim, _, dtype, mode, eb = case
    x, dims = _input(case)
    n = x.size
    want = oracle.compress(x, eb, mode, dim, dims)
    print(f"scan stress {_ids(case)} {knobs} n={n}", flush=True)
    stress(**knobs)
    got = gpu_compress(x, eb, mode, dim, dims)
    rc = P.scan_report()
    assert got == want
    assert rc["stolen"] and rc["claims"] > 0, rc
    # these streams are a single segment of the decompression's offset pass
    # (32 KiB of header bytes): nothing to claim, the result must not change
    y = gpu_decompress(want, n, dtype, eb, mode, dim, dims)
    yref = oracle.decompress(want, n, dtype, eb, mode, dim, dims)
    assert np.array_equal(y.view(np.uint8), yref.view(np.uint8))


# Streams of 10-12 segments of the decompression's offset pass (a segment is
# 32,768 header bytes, i.e. 32,768 blocks), so that held segments have
# successors that claim them: (dim, dims, dtype, mode, eb)
OFFSET_CASES = [
    (1, None, np.float32, 1, 1e-3),                # 10,485,767 values: 10.0 segments
    (1, None, np.float64, 2, 1e-4),
    (2, (2048, 1160, 8), np.float32, 2, 1e-2),     # 256 x 145 x 8 = 296,960 blocks
    (3, (256, 256, 300), np.float64, 0, 1e-3),     # 64 x 64 x 75 = 307,200 blocks
]


def _offset_input(case):
    dim, dims, dtype, mode, eb = case
    if dim == 1:
        return mixed_field(10 * 1024 * 1024 + 7, seed=9, dtype=dtype), (0, 0, 0)
    x, y, z = dims
    i = np.arange(x * y * z, dtype=np.float64)
    rng = np.random.default_rng(10)
    v = np.sin(i * 3e-5) * 40 + np.sin(i * 7e-3) * 3 + rng.normal(0, 0.05, i.size)
    return v.astype(dtype), dims


@pytest.mark.parametrize("knobs", [HOLD, AT_ONCE], ids=["hold", "steal-at-once"])
@pytest.mark.parametrize("case", OFFSET_CASES, ids=[_ids(c) for c in OFFSET_CASES])
def test_offset_pass_claims(device, oracle, stress, case, knobs):
    """Single-array decompression: its offset pass (the scan of the header's
    block sizes) under the same stress; the decompressions equal the oracle's
    and claims are reported."""
    from cuszp_amd import pycuSZp as P

    dim, _, dtype, mode, eb = case
    x, dims = _offset_input(case)
    n = x.size
    want = oracle.compress(x, eb, mode, dim, dims, nthreads=oracle.max_threads())
    print(f"offset pass stress {_ids(case)} {knobs} n={n}", flush=True)
    stress(**knobs)
    y = gpu_decompress(want, n, dtype, eb, mode, dim, dims)
    rd = P.scan_report()
    yref = oracle.decompress(want, n, dtype, eb, mode, dim, dims, nthreads=oracle.max_threads())
    assert np.array_equal(y.view(np.uint8), yref.view(np.uint8))
    assert rd["stolen"] and rd["claims"] > 0, rd


@pytest.mark.parametrize("knobs", [HOLD, AT_ONCE], ids=["hold", "steal-at-once"])
def test_batched_claims(device, oracle, stress, knobs):
    """A mixed batch (every array its own scan segment) under the same stress:
    streams and decompressions equal the oracle's, claims are reported."""
    import torch
    from cuszp_amd import pycuSZp as P

    arrays = []
    for k, case in enumerate(CASES):
        x, dims = _input(case)
        arrays.append((x, dims, case))
    # and a few small 1D arrays between them (single-tile segments)
    for k in range(3):
        x = mixed_field(5000 + 777 * k, seed=k, dtype=np.float32)
        arrays.append((x, (0, 0, 0), (1, None, np.float32, k, 1e-3)))
    stress(**knobs)
    ins, outs, citems = [], [], []
    for x, dims, (dim, _, dtype, mode, eb) in arrays:
        t = torch.from_numpy(x).to(device)
        cap = P.max_compressed_size(x.size, dim, dims, 1 if dtype == np.float64 else 0, mode)
        o = torch.zeros(cap, dtype=torch.uint8, device=device)
        ins.append(t)
        outs.append(o)
        citems.append(P.batch_item(t, o, x.size, cap, eb, dim, dims, 1 if dtype == np.float64 else 0, mode))
    print(f"batched scan stress {knobs}", flush=True)
    with P.BatchPlan(citems, P.BATCH_COMPRESS) as plan:
        plan.execute()
        sizes, st = plan.results()
        rc = plan.scan_report()
    assert all(v == 0 for v in st), st
    assert rc["stolen"] and rc["claims"] > 0, rc
    streams = []
    for (x, dims, (dim, _, dtype, mode, eb)), o, s in zip(arrays, outs, sizes):
        want = oracle.compress(x, eb, mode, dim, dims)
        got = o[:s].cpu().numpy().tobytes()
        assert got == want
        streams.append(want)
    ys, ditems = [], []
    for (x, dims, (dim, _, dtype, mode, eb)), o, s in zip(arrays, outs, sizes):
        y = torch.full((x.size,), float("nan"), dtype=torch.float64 if dtype == np.float64 else torch.float32,
                       device=device)
        ys.append(y)
        ditems.append(P.batch_item(o, y, x.size, s, eb, dim, dims, 1 if dtype == np.float64 else 0, mode))
    with P.BatchPlan(ditems, P.BATCH_DECOMPRESS) as plan:
        plan.execute()
        _, st = plan.results()
        rd = plan.scan_report()
    assert all(v == 0 for v in st), st
    assert rd["stolen"] and rd["claims"] > 0, rd
    for (x, dims, (dim, _, dtype, mode, eb)), y, s in zip(arrays, ys, streams):
        yref = oracle.decompress(s, x.size, dtype, eb, mode, dim, dims)
        assert np.array_equal(y.cpu().numpy().view(np.uint8), yref.view(np.uint8))


def test_report_without_stress(device, oracle, stress):
    """Normal settings with the counters on: the stream is unchanged, the
    counters are consistent (every waiting lookback polled at least once), and
    the report clears after reading."""
    from cuszp_amd import pycuSZp as P

    stress(stats=True)
    x = mixed_field(2_000_003, seed=3, dtype=np.float32)
    want = oracle.compress(x, 1e-3, 1)
    assert gpu_compress(x, 1e-3, 1) == want
    r = P.scan_report()
    assert r["polls"] >= r["waits"] >= 0
    assert r["claims"] == 0 or r["stolen"]
    again = P.scan_report()
    assert again["polls"] == 0 and again["claims"] == 0 and not again["stolen"]


def test_scan_config_rejects_long_holds(device):
    from cuszp_amd import pycuSZp as P

    with pytest.raises(P.CuSZpError):
        P.scan_config(hold_every=2, hold_ticks=20_000_000)
    P.scan_config()
