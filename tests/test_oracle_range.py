"""The oracle's relative error bound (examples/cuSZp.cpp:176-210 of the
reference), pinned against numpy: the CLI's comparison loop equals
nanmin/nanmax whenever element 0 is not NaN, the bound is the float product
rel * (max - min) (f64: the difference rounded to float first), and a NaN in
element 0 makes everything NaN."""
import numpy as np
import pytest


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("seed", range(4))
def test_rel_eb_matches_numpy(oracle, dtype, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 50_000))
    x = (rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 3)).astype(dtype)
    if n > 10:
        x[rng.integers(1, n, 5)] = np.nan   # never element 0
    eb, mn, mx = oracle.rel_eb(x, 1e-3)
    assert mn == np.nanmin(x) and mx == np.nanmax(x)
    if dtype == np.float32:
        want = np.float32(1e-3) * (np.float32(mx) - np.float32(mn))
    else:
        want = np.float32(1e-3) * np.float32(np.float64(mx) - np.float64(mn))
    assert np.float32(eb) == want


def test_rel_eb_special_values(oracle):
    x = np.array([2.0, -np.inf, 1.0], np.float32)
    assert oracle.rel_eb(x, 0.5)[1:] == (-np.inf, 2.0)
    x = np.array([np.nan, 1.0, 2.0], np.float32)
    eb, mn, mx = oracle.rel_eb(x, 0.5)
    assert np.isnan(eb) and np.isnan(mn) and np.isnan(mx)
    x = np.array([7.0], np.float64)
    assert oracle.rel_eb(x, 0.5) == (0.0, 7.0, 7.0)
    # the f32 difference is taken in float: 16777217 is not representable
    x = np.array([0.0, 16777216.0, -1.0], np.float32)
    eb, mn, mx = oracle.rel_eb(x, 1.0)
    assert eb == float(np.float32(16777216.0) - np.float32(-1.0)) == 16777216.0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_quality_matches_numpy(oracle, dtype):
    """computePSNR (src/cuSZp_utility.cu:891-950) against the textbook formulas."""
    rng = np.random.default_rng(7)
    x = (rng.standard_normal(200_001) * 3).astype(dtype)
    x[::97] = 0
    y = (x + rng.uniform(-1e-2, 1e-2, x.size)).astype(dtype)
    q = oracle.quality(x, y)
    xd, yd = x.astype(np.float64), y.astype(np.float64)
    err = np.abs(yd - xd)
    mse = np.mean(err ** 2)
    rng_ = xd.max() - xd.min()
    assert q["range"] == rng_
    assert q["max_abs_err"] == err.max()
    nz = x != 0
    assert q["max_rel_err"] == pytest.approx(np.max(err[nz] / np.abs(xd[nz])), rel=1e-6)
    assert q["mse"] == pytest.approx(mse, rel=1e-6)
    assert q["psnr"] == pytest.approx(20 * np.log10(rng_) - 10 * np.log10(mse), rel=1e-6)
    assert q["norm_err"] == pytest.approx(np.sqrt(np.sum(err ** 2)), rel=1e-6)
    assert q["norm_err_norm"] == pytest.approx(np.sqrt(np.sum(err ** 2)) / np.sqrt(np.sum(yd ** 2)), rel=1e-6)
    assert q["nrmse"] == pytest.approx(np.sqrt(mse) / rng_, rel=1e-6)
    assert q["correlation"] == pytest.approx(np.corrcoef(xd, yd)[0, 1], rel=1e-9)
    assert oracle.quality(x, x)["max_abs_err"] == 0.0


def _ssim_numpy(x, y, dims):
    """Windowed SSIM restated with numpy's sliding windows (float64 statistics;
    the window range in the values' type)."""
    X, Y, Z = dims
    a = x.reshape(Z, Y, X)
    b = y.reshape(Z, Y, X)
    wa = np.lib.stride_tricks.sliding_window_view(a, (7, 7, 7))[::2, ::2, ::2]
    wb = np.lib.stride_tricks.sliding_window_view(b, (7, 7, 7))[::2, ::2, ::2]
    wa64, wb64 = wa.astype(np.float64), wb.astype(np.float64)
    mx, my = wa64.mean(axis=(3, 4, 5)), wb64.mean(axis=(3, 4, 5))
    vx = ((wa64 - mx[..., None, None, None]) ** 2).mean(axis=(3, 4, 5))
    vy = ((wb64 - my[..., None, None, None]) ** 2).mean(axis=(3, 4, 5))
    vxy = ((wa64 - mx[..., None, None, None]) * (wb64 - my[..., None, None, None])).mean(axis=(3, 4, 5))
    r = (wa.max(axis=(3, 4, 5)) - wa.min(axis=(3, 4, 5))).astype(np.float64)
    c1 = np.where(r == 0, 1e-4, 1e-4 * r * r)
    c2 = np.where(r == 0, 9e-4, 9e-4 * r * r)
    sx, sy = np.sqrt(vx), np.sqrt(vy)
    s = (2 * mx * my + c1) / (mx * mx + my * my + c1) * (2 * sx * sy + c2) / (sx * sx + sy * sy + c2) \
        * (vxy + c2 / 2) / (sx * sy + c2 / 2)
    return float(s.mean())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ssim_matches_numpy(oracle, dtype):
    """SSIM (src/cuSZp_utility.cu:761-889): 1 for identical volumes, and the
    windowed formulas restated with numpy agree."""
    rng = np.random.default_rng(5)
    dims = (20, 15, 11)
    x = np.cumsum(rng.standard_normal(20 * 15 * 11)).astype(dtype)
    y = (x + rng.uniform(-0.3, 0.3, x.size)).astype(dtype)
    assert oracle.ssim(x, x, dims) == pytest.approx(1.0, abs=1e-12)
    assert oracle.ssim(x, y, dims) == pytest.approx(_ssim_numpy(x, y, dims), rel=1e-9)
    flat = np.full(x.size, 2.5, dtype)   # zero range: the constants fall back to (0.01)^2, (0.03)^2
    assert oracle.ssim(flat, flat, dims) == pytest.approx(1.0, abs=1e-12)
    assert oracle.ssim(x[:6 * 15 * 11], y[:6 * 15 * 11], (6, 15, 11)) < -1   # a window does not fit: refused
