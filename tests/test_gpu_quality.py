"""Reconstruction quality on the device (cuSZp_quality) against the oracle's
sequential restatement of src/cuSZp_utility.cu:891-950: maxima and the value
range exactly, the double sums (summed in another order) to 1e-9 relative,
on a real decompression of both types."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 1000, 1_048_577])
def test_quality_matches_oracle(device, oracle, dtype, n):
    import torch
    from cuszp_amd import pycuSZp as P

    rng = np.random.default_rng(n)
    x = np.cumsum(rng.standard_normal(n)).astype(dtype)
    x[::13] = 0
    t = torch.from_numpy(x).to(device)
    eb = 1e-2
    cmp, size = P.compress_tensor(t, eb, mode="outlier")
    y = P.decompress_tensor(cmp, size, n, t.dtype, eb, mode="outlier")
    got = P.quality(t, y)
    want = oracle.quality(x, y.cpu().numpy())
    for k in ("range", "max_abs_err", "max_rel_err"):
        assert got[k] == want[k], k
    for k in ("mse", "psnr", "norm_err", "norm_err_norm", "nrmse", "correlation"):
        if np.isfinite(want[k]):
            assert got[k] == pytest.approx(want[k], rel=1e-9, abs=1e-300), k
        else:
            assert not np.isfinite(got[k]) or n == 1, k
    assert got["max_abs_err"] <= eb * (1 + 2 ** -20) + 4 * np.finfo(dtype).eps * float(np.abs(x).max())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ssim_matches_oracle(device, oracle, dtype):
    import torch
    from cuszp_amd import pycuSZp as P
    from tests.workloads import grf_volume

    dims = (40, 36, 30)
    x = grf_volume(30, 36, 40, seed=6, dtype=dtype).ravel()
    t = torch.from_numpy(x).to(device)
    cmp, size = P.compress_tensor(t, 0.5, mode="outlier", dim=3, dims=dims)
    y = P.decompress_tensor(cmp, size, t.numel(), t.dtype, 0.5, mode="outlier", dim=3, dims=dims)
    got = P.ssim(t, y, dims)
    want = oracle.ssim(x, y.cpu().numpy(), dims)
    assert got == pytest.approx(want, rel=1e-9)
    assert P.ssim(t, t, dims) == pytest.approx(1.0, abs=1e-12)
    with pytest.raises(P.CuSZpError):
        P.ssim(t[: 6 * 36 * 30].contiguous(), y[: 6 * 36 * 30].contiguous(), (6, 36, 30))
