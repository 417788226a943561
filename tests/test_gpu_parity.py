"""Parity of the HIP kernels (through the C ABI) with the CPU oracle.

For every (dim, mode, dtype) and a spread of shapes, including ragged edges:
  * the compressed stream is byte-identical to the oracle's,
  * the decompressed array is bit-identical to the oracle's decompression,
  * the error bound holds (tolerance in tests/helpers.py).
"""
import numpy as np
import pytest

from tests.helpers import check_error_bound, gpu_compress, gpu_decompress
from tests.workloads import mixed_field

pytestmark = pytest.mark.gpu

SHAPES_1D = [1, 7, 31, 32, 33, 63, 64, 65, 1000, 32767, 32768, 32769, 65536 + 5, 300_001]
SHAPES_2D = [(4, 1, 1), (8, 8, 1), (12, 10, 3), (4, 17, 2), (64, 64, 4), (516, 7, 1), (1024, 130, 3)]
SHAPES_3D = [(4, 4, 4), (8, 9, 5), (12, 3, 2), (4, 1, 1), (64, 64, 64), (100, 36, 17)]

CASES = [(1, (0, 0, 0), n) for n in SHAPES_1D]
CASES += [(2, d, d[0] * d[1] * d[2]) for d in SHAPES_2D]
CASES += [(3, d, d[0] * d[1] * d[2]) for d in SHAPES_3D]


def _id(c):
    dim, dims, n = c
    return f"{dim}D-{n}" if dim == 1 else f"{dim}D-{dims[0]}x{dims[1]}x{dims[2]}"


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("mode", [0, 1, 2], ids=["fixed", "plain", "outlier"])
@pytest.mark.parametrize("case", CASES, ids=[_id(c) for c in CASES])
def test_parity(device, oracle, case, mode, dtype):
    dim, dims, n = case
    x = mixed_field(n, seed=n * 7 + mode, dtype=dtype)
    eb = 1e-2
    ebT = float(dtype(eb))
    ref = oracle.compress(x, eb, mode, dim, dims)
    got = gpu_compress(x, eb, mode, dim, dims)
    assert len(got) == len(ref)
    if got != ref:
        a = np.frombuffer(got, np.uint8)
        b = np.frombuffer(ref, np.uint8)
        i = int(np.argmax(a != b))
        pytest.fail(f"stream differs at byte {i} of {len(ref)} (header {oracle.header_size(dim, n, dims)})")
    dref = oracle.decompress(ref, n, dtype, eb, mode, dim, dims)
    dgot = gpu_decompress(ref, n, dtype, eb, mode, dim, dims)
    assert np.array_equal(dgot.view(np.uint8), dref.view(np.uint8))
    check_error_bound(x, dgot, ebT)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("mode", [0, 1, 2], ids=["fixed", "plain", "outlier"])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_special_values(device, oracle, dim, mode, dtype):
    """NaN, +-inf, +-0, subnormals, exact bin ties and saturating magnitudes."""
    rng = np.random.default_rng(5 + dim)
    dims = (0, 0, 0) if dim == 1 else ((16, 8, 4) if dim == 2 else (8, 8, 8))
    n = 4096 if dim == 1 else 512
    eb = 0.5  # r = 1: quantisation code = round(x)
    x = rng.normal(0, 4, n)
    specials = [np.nan, np.inf, -np.inf, 0.0, -0.0, 0.5, -0.5, 1.5, -1.5, 2.5, -2.5,
                2.0**31, -(2.0**31), 3e9, -3e9, 1e300 if dtype == np.float64 else 3e38,
                -2147483647.5, -2147483648.5, 2147483647.5, 5e-324 if dtype == np.float64 else 1e-45,
                np.finfo(dtype).tiny, -np.finfo(dtype).tiny, 1e-40]
    idx = rng.choice(n, len(specials) * 4, replace=False)
    for k, i in enumerate(idx):
        x[i] = specials[k % len(specials)]
    x = x.astype(dtype)
    ref = oracle.compress(x, eb, mode, dim, dims)
    got = gpu_compress(x, eb, mode, dim, dims)
    assert got == ref
    dref = oracle.decompress(ref, n, dtype, eb, mode, dim, dims)
    dgot = gpu_decompress(ref, n, dtype, eb, mode, dim, dims)
    assert np.array_equal(dgot.view(np.uint8), dref.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("mode", [0, 1, 2], ids=["fixed", "plain", "outlier"])
@pytest.mark.parametrize("dim", [2, 3])
def test_residual_int_min(device, oracle, dim, mode, dtype):
    """A residual of -2^31 gives rate 32 in 2D/3D (unsigned max in the reference)."""
    dims = (8, 8, 1) if dim == 2 else (4, 4, 4)
    n = 64
    x = np.zeros(n, np.float64)
    x[0] = -1.0
    x[1] = 3e9             # code INT_MAX after -1 -> residual wraps to INT_MIN (plain/outlier)
    x[9] = -2147483647.5   # code INT_MIN in f64 (fixed mode)
    x[20] = 5.0
    x = x.astype(dtype)
    ref = oracle.compress(x, 0.5, mode, dim, dims)
    import warnings

    from cuszp_amd.pycuSZp import BoundExceededWarning

    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        got = gpu_compress(x, 0.5, mode, dim, dims)
    assert got == ref
    lossy = [w for w in caught if issubclass(w.category, BoundExceededWarning)]
    # outlier mode keeps R1 = 32 as the plain byte 32, which decoders read as
    # rate 0 (the block decodes to zeros): the call warns; fixed and plain
    # mode encode all 32 planes and decode them exactly
    assert len(lossy) == (1 if mode == 2 else 0)
    if mode == 2:
        assert ref[0] == 32   # header byte of the block: plain rate 32
    dref = oracle.decompress(ref, n, dtype, 0.5, mode, dim, dims)
    dgot = gpu_decompress(ref, n, dtype, 0.5, mode, dim, dims)
    assert np.array_equal(dgot.view(np.uint8), dref.view(np.uint8))
    if mode == 2:
        assert not dgot.any()   # the whole block decoded as zeros


@pytest.mark.parametrize("eb", [1e-1, 1e-3, 1e-6])
@pytest.mark.parametrize("mode", [0, 1, 2], ids=["fixed", "plain", "outlier"])
def test_error_bounds_1d(device, oracle, mode, eb):
    x = mixed_field(200_000, seed=11, dtype=np.float32)
    ref = oracle.compress(x, eb, mode)
    got = gpu_compress(x, eb, mode)
    assert got == ref
    dgot = gpu_decompress(got, x.size, np.float32, eb, mode)
    check_error_bound(x, dgot, float(np.float32(eb)))


def test_rate_zero_blocks_written(device, oracle):
    """Rate-0 blocks decode to +0.0 even into a dirty output buffer."""
    x = np.zeros(70_000, np.float32)
    x[40_000:40_010] = 3.0
    got = gpu_compress(x, 1e-2, 1)
    dgot = gpu_decompress(got, x.size, np.float32, 1e-2, 1, fill=7.0)
    dref = oracle.decompress(got, x.size, np.float32, 1e-2, 1)
    assert np.array_equal(dgot.view(np.uint32), dref.view(np.uint32))
    assert np.all(dgot[:40_000] == 0) and not np.signbit(dgot[:40_000]).any()


# error bounds at the edges of the type's range: smallest normal, subnormal
# (0.5 / eb overflows to inf, so every nonzero value saturates and 0 * inf is
# NaN), largest whose double is finite, and beyond
EB_EDGES = [
    (np.float32, 2.0**-126), (np.float32, 2.0**125), (np.float32, 1e-40), (np.float32, 3e38),
    (np.float32, 2.0**-126 * (1 - 2.0**-23)), (np.float32, 2.0**126),
    (np.float64, 2.0**-1022), (np.float64, 2.0**1021), (np.float64, 1e-310), (np.float64, 1.5e308),
]


@pytest.mark.parametrize("dim", [1, 3], ids=["1D", "3D"])
@pytest.mark.parametrize("case", EB_EDGES, ids=[f"{np.dtype(d).name}-{e:.3g}" for d, e in EB_EDGES])
def test_error_bound_range_edges(oracle, dim, case):
    """Streams byte-identical and decompressions bit-identical to the oracle
    for error bounds at the edges of the type's range, in arrays of whole
    tiles plus a ragged tail (1D) or a partial last z-layer of blocks (3D).
    The padding elements of the ragged tiles must code as q = 0 even when
    0.5 / eb is inf (src/cuSZp_kernels_1D_f32.cu:113-134 zero them)."""
    dtype, eb = case
    rng = np.random.default_rng(17)
    if dim == 1:
        n, dims = 4 * 16384 + 100, (0, 0, 0)
    else:
        dims = (64, 32, 33)
        n = dims[0] * dims[1] * dims[2]
    fmax = float(np.finfo(dtype).max)
    # values of a few hundred bins where eb allows it, else as large as the type allows
    amp = min(eb * 600.0, fmax / 2)
    x = (np.sin(np.arange(n) * 0.01) * amp + rng.standard_normal(n) * min(eb * 3.0, fmax / 8)).astype(dtype)
    x[::977] = 0
    for mode in (0, 1, 2):
        got = gpu_compress(x, eb, mode, dim, dims)
        want = oracle.compress(x, eb, mode, dim, dims)
        assert got == want, (mode, len(got), len(want))
        y = gpu_decompress(got, n, dtype, eb, mode, dim, dims)
        yref = oracle.decompress(want, n, dtype, eb, mode, dim, dims)
        assert np.array_equal(y.view(np.uint8), yref.view(np.uint8)), mode
