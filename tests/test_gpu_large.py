"""Large arrays: the 8-wave tiles the library uses from 128 tiles up, against
the oracle (byte-identical streams, bit-identical decompression), and the
size-independent properties at the benchmark sizes (error bound, round trip
through the stream, re-compression of the decompressed array, batched = single).
"""
import numpy as np
import pytest

from tests.helpers import check_error_bound, gpu_compress, gpu_decompress
from tests.workloads import grf_volume, sinusoid_field, weights

pytestmark = pytest.mark.gpu

# (dim, dims, generator, dtype, mode, eb): each at least 128 8-wave tiles
LARGE = [
    (1, (0, 0, 0), "sin", np.float32, 1, 1e-3),          # 8 Mi values
    (1, (0, 0, 0), "sin", np.float64, 2, 1e-4),          # 4 Mi values
    (2, (2048, 2048, 1), "w", np.float64, 0, 1e-3),      # fixed mode, weight-like
    (2, (1024, 512, 9), "grf", np.float32, 2, 1e-2),     # 2D over 9 slices
    (3, (256, 256, 128), "grf", np.float32, 2, 5e-3),    # 3D outlier
    (3, (128, 128, 192), "grf", np.float64, 1, 1e-3),    # 3D plain f64
]


def _field(kind, dims, n, dtype, seed):
    if kind == "sin":
        return sinusoid_field(n, seed=seed, zero_runs=True, dtype=dtype)
    if kind == "w":
        return weights(dims[1], dims[0], 0.02, seed=seed, dtype=dtype).ravel()
    x, y, z = dims
    return grf_volume(z, y, x, seed=seed, dtype=dtype).ravel()


@pytest.mark.parametrize("case", LARGE, ids=[f"{c[0]}D-{c[2]}-{np.dtype(c[3]).name}-m{c[4]}" for c in LARGE])
def test_large_parity(device, oracle, case):
    dim, dims, kind, dtype, mode, eb = case
    n = (8 << 20 if dtype == np.float32 else 4 << 20) if dim == 1 else dims[0] * dims[1] * dims[2]
    x = _field(kind, dims, n, dtype, seed=n % 97)
    print(f"case {case} n={n}", flush=True)
    got = gpu_compress(x, eb, mode, dim, dims)
    want = oracle.compress(x, eb, mode, dim, dims, nthreads=oracle.max_threads())
    assert len(got) == len(want)
    assert got == want
    y = gpu_decompress(got, n, dtype, eb, mode, dim, dims)
    yref = oracle.decompress(want, n, dtype, eb, mode, dim, dims, nthreads=oracle.max_threads())
    assert np.array_equal(y.view(np.uint8), yref.view(np.uint8))
    check_error_bound(x, y, eb)


def test_full_size_round_trip_properties(device):
    """2^28 f32 (1 GiB) in plain mode: error bound on the device, the stream
    of the decompressed array equals the original stream (re-compression is
    exact for these codes), and the batched form gives the same bytes."""
    import torch
    from cuszp_amd import pycuSZp as P

    n = 1 << 28
    g = torch.Generator(device=device).manual_seed(11)
    i = torch.arange(n, dtype=torch.float64, device=device)
    x = (torch.sin(i * 3e-5) + 0.5 * torch.sin(i * 7e-4 + 1.0)).float()
    x += torch.randn(n, device=device, generator=g) * 1e-3
    del i
    eb = P.rel_to_abs(x, 1e-3)
    c, size = P.compress_tensor(x, eb, mode="plain")
    y = P.decompress_tensor(c, size, n, torch.float32, eb, mode="plain")
    ebf = float(np.float32(eb))
    err = (x.double() - y.double()).abs()
    tol = ebf * (1 + 2.0**-20) + (x.double().abs() + y.double().abs()) * 2 * 2.0**-24
    assert bool((err <= tol).all()), float((err / ebf).max())
    c2, size2 = P.compress_tensor(y, eb, mode="plain")
    assert size2 == size and torch.equal(c2[:size], c[:size])
    (cb, sb), = P.compress_tensors([x], [eb], ["plain"])
    assert sb == size and torch.equal(cb[:size], c[:size])
    q = P.quality(x, y)
    assert q["max_abs_err"] <= ebf * (1 + 2.0**-20) + 4 * 2.0**-24 * float(x.abs().max())


# more than 2^31 elements: every element and stream offset needs 64 bits
BEYOND_2_31 = [
    (1, (0, 0, 0), (1 << 31) + 4133, "outlier", 1e-3),     # 8 GiB of f32, ragged last block
    (3, (2048, 1024, 1025), 2048 * 1024 * 1025, "plain", 1e-3),   # z not a multiple of 4
]


@pytest.mark.parametrize("case", BEYOND_2_31, ids=["1D-outlier-2^31+4133", "3D-plain-2048x1024x1025"])
def test_beyond_2_31_elements_parity(device, oracle, case):
    """Byte-identical stream and bit-identical decompression against the
    oracle for f32 arrays of more than 2^31 elements (generated on the device)."""
    import torch
    from cuszp_amd import pycuSZp as P

    dim, dims, n, mode, eb = case
    print(f"case {case}: generating", flush=True)
    if dim == 1:
        g = torch.Generator(device=device).manual_seed(5)
        x = torch.sin(torch.arange(n, dtype=torch.float32, device=device) * 1e-4)
        x += torch.randn(n, device=device, generator=g) * 2e-3
    else:
        X, Y, Z = dims
        ax = torch.sin(torch.arange(X, dtype=torch.float32, device=device) * 0.013)
        ay = torch.cos(torch.arange(Y, dtype=torch.float32, device=device) * 0.007)
        az = torch.sin(torch.arange(Z, dtype=torch.float32, device=device) * 0.011) * 0.5
        x = (az[:, None, None] + ay[None, :, None] + ax[None, None, :]).reshape(-1).contiguous()
    assert x.numel() == n and n > (1 << 31)
    c, size = P.compress_tensor(x, eb, mode=mode, dim=dim, dims=dims if dim != 1 else None)
    y = P.decompress_tensor(c, size, n, torch.float32, eb, mode=mode, dim=dim, dims=dims if dim != 1 else None)
    ebf = float(np.float32(eb))
    worst = 0.0
    for s in range(0, n, 1 << 28):   # error bound on the device, in slices
        e = (x[s:s + (1 << 28)].double() - y[s:s + (1 << 28)].double()).abs()
        worst = max(worst, float(e.max()))
    assert worst <= ebf * (1 + 2.0**-20) + 4 * 2.0**-24 * float(x.abs().max())
    torch.cuda.synchronize()
    got = c[:size].cpu().numpy().tobytes()
    del c
    print(f"case {case}: stream {size} bytes; oracle compress", flush=True)
    xh = x.cpu().numpy()
    del x
    want = oracle.compress(xh, eb, P.MODES[mode], dim, dims, nthreads=oracle.max_threads())
    assert len(got) == len(want)
    assert got == want
    del want
    print(f"case {case}: oracle decompress", flush=True)
    yref = oracle.decompress(got, n, np.float32, eb, P.MODES[mode], dim, dims, nthreads=oracle.max_threads())
    yh = y.cpu().numpy()
    del y
    assert np.array_equal(yh.view(np.uint32), yref.view(np.uint32))
