"""The C++ oracle against the independent numpy restatement, plus round-trip
properties of the format on small random inputs (all dims, modes, dtypes)."""
import numpy as np
import pytest

from oracle import np_oracle as NPO
from tests.helpers import check_error_bound
from tests.workloads import mixed_field

CASES = [(1, (0, 0, 0), n) for n in (1, 31, 32, 33, 1000, 32769)]
CASES += [(2, d, d[0] * d[1] * d[2]) for d in [(4, 1, 1), (12, 10, 3), (4, 17, 2), (20, 9, 1)]]
CASES += [(3, d, d[0] * d[1] * d[2]) for d in [(4, 4, 4), (8, 9, 5), (12, 3, 2), (4, 1, 7)]]


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("mode", [0, 1, 2], ids=["fixed", "plain", "outlier"])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}D-{c[2]}" for c in CASES])
def test_cpp_equals_numpy(oracle, case, mode, dtype):
    dim, dims, n = case
    x = mixed_field(n, seed=n + 3 * mode, dtype=dtype)
    eb = 1e-2
    c1 = oracle.compress(x, eb, mode, dim, dims)
    assert c1 == NPO.compress(x, eb, mode, dim, dims)
    assert len(c1) >= oracle.header_size(dim, n, dims)
    assert len(c1) <= oracle.max_size(dim, n, dims)
    y1 = oracle.decompress(c1, n, dtype, eb, mode, dim, dims)
    y2 = NPO.decompress(c1, n, dtype, eb, mode, dim, dims)
    assert np.array_equal(y1.view(np.uint8), y2.view(np.uint8))
    check_error_bound(x, y1, float(dtype(eb)))


def test_idempotent_recompression(oracle):
    """Decompressed values sit on bin centres: compressing them again (same eb)
    reproduces the same codes and the same stream."""
    x = mixed_field(50_000, seed=9)
    c = oracle.compress(x, 1e-3, 1)
    y = oracle.decompress(c, x.size, np.float32, 1e-3, 1)
    assert oracle.compress(y, 1e-3, 1) == c


def test_truncated_stream_rejected(oracle):
    x = mixed_field(70_000, seed=4)
    c = oracle.compress(x, 1e-3, 1)
    with pytest.raises(ValueError):
        oracle.decompress(c[:-1], x.size, np.float32, 1e-3, 1)
    with pytest.raises(ValueError):
        oracle.decompress(c[:100], x.size, np.float32, 1e-3, 1)


def test_sawtooth_reference_driver(oracle):
    """The reference test drivers' data: sawtooth -20..20 step 0.1, eb = 1e-2 * range,
    checked at |x - x'| <= 1.1 eb (examples/cuSZp_test_f32_1D.cpp:27-52,108)."""
    n = 1 << 16
    x = np.empty(n, np.float32)
    v, step = np.float32(-20.0), np.float32(0.1)
    for i in range(n):       # float accumulation as in the driver
        x[i] = v
        v = np.float32(v + step)
        if v > np.float32(20.0):
            v = np.float32(-20.0)
    eb = float((x.max() - x.min()) * np.float32(1e-2))
    for mode in (0, 1, 2):
        for dim, dims in [(1, (0, 0, 0)), (2, (256, 256, 1)), (3, (64, 32, 32))]:
            c = oracle.compress(x, eb, mode, dim, dims)
            y = oracle.decompress(c, n, np.float32, eb, mode, dim, dims)
            assert np.max(np.abs(x.astype(np.float64) - y)) <= 1.1 * eb
