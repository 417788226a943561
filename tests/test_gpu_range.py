"""Relative error bounds on the device (cuSZp_value_range / cuSZp_rel_error_bound)
against the oracle's restatement of the reference CLI (examples/cuSZp.cpp:176-210):
the same min, max and float bound, for every alignment of the array start,
NaNs anywhere, infinities, single elements and both types."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(P, oracle, torch, device, host, rel=1e-3):
    t = torch.from_numpy(host).to(device)
    eb_ref, mn_ref, mx_ref = oracle.rel_eb(host, rel)
    mn, mx = P.value_range(t)
    eb = P.rel_to_abs(t, rel)
    if np.isnan(mn_ref):
        assert np.isnan(mn) and np.isnan(mx) and np.isnan(eb)
        return
    assert (mn, mx) == (mn_ref, mx_ref)
    assert np.float32(eb) == np.float32(eb_ref) or (np.isnan(eb) and np.isnan(eb_ref))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 3, 17, 1000, 65_537, 3_000_001])
def test_range_matches_oracle(device, oracle, dtype, n):
    import torch
    from cuszp_amd import pycuSZp as P

    rng = np.random.default_rng(n)
    base = (rng.standard_normal(n + 8) * 100).astype(dtype)
    for off in range(4):   # array starts at every 4/8-byte offset of a 16-byte line
        _check(P, oracle, torch, device, np.ascontiguousarray(base[off:off + n]))
    x = base[:n].copy()
    if n > 2:
        x[rng.integers(1, n, 3)] = np.nan
        x[n // 2] = 1e30
        _check(P, oracle, torch, device, x)
        x[-1] = -np.inf
        _check(P, oracle, torch, device, x)
    x[0] = np.nan   # the CLI's loop starts from element 0: everything NaN
    _check(P, oracle, torch, device, x)


def test_range_repeats_and_streams(device, oracle):
    import torch
    from cuszp_amd import pycuSZp as P

    host = np.random.default_rng(5).uniform(-3, 9, 1 << 22).astype(np.float32)
    t = torch.from_numpy(host).to(device)
    _, mn_ref, mx_ref = oracle.rel_eb(host, 1e-2)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()   # t is complete before another stream reads it
    for k in range(4):   # the per-stream completion ticket is reset by every launch
        assert P.value_range(t, stream=s if k % 2 else None) == (mn_ref, mx_ref)
    out, n, eb = P.compress_tensor_rel(t, 1e-2, mode="plain")
    assert eb == oracle.rel_eb(host, 1e-2)[0]
    c = oracle.compress(host, eb, 1)
    assert out[:n].cpu().numpy().tobytes() == c
