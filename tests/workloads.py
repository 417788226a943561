"""Seeded synthetic inputs of the benchmark configurations (numpy, CPU).

sinusoid_field   8 sinusoids + N(0,1e-3) noise, optional 20 % zero runs
grf_volume       Gaussian random field, k^(-11/3) power spectrum, [-50,50], 0.1 % spikes
weights          N(0, sigma) dense-layer-like matrix
"""
from __future__ import annotations

import numpy as np


def sinusoid_field(n: int, seed: int = 0, zero_runs: bool = True, dtype=np.float32) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    f = rng.uniform(1e-6, 1e-3, 8)
    a = rng.uniform(0.1, 1.0, 8)
    ph = rng.uniform(0, 2 * np.pi, 8)
    x = np.zeros(n, np.float64)
    chunk = 1 << 22
    for s in range(0, n, chunk):
        i = np.arange(s, min(n, s + chunk), dtype=np.float64)
        v = np.zeros(i.size)
        for k in range(8):
            v += a[k] * np.sin(2 * np.pi * f[k] * i + ph[k])
        x[s:s + i.size] = v
    x += rng.normal(0.0, 1e-3, n)
    if zero_runs:
        target = int(0.2 * n)
        covered = np.zeros(n, bool)
        while covered.sum() < target:
            ln = int(rng.uniform(32, 4096))
            st = int(rng.integers(0, max(1, n - ln)))
            covered[st:st + ln] = True
        x[covered] = 0.0
    return x.astype(dtype)


def grf_volume(nz: int, ny: int, nx: int, seed: int = 1, dtype=np.float32, spikes: float = 1e-3) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.standard_normal((nz, ny, nx))
    F = np.fft.rfftn(w)
    kz = np.fft.fftfreq(nz)[:, None, None]
    ky = np.fft.fftfreq(ny)[None, :, None]
    kx = np.fft.rfftfreq(nx)[None, None, :]
    k = np.sqrt(kz**2 + ky**2 + kx**2)
    k[0, 0, 0] = 1.0
    F *= k ** (-11.0 / 6.0)
    F[0, 0, 0] = 0.0
    v = np.fft.irfftn(F, s=(nz, ny, nx))
    v = (v - v.min()) / (v.max() - v.min()) * 100.0 - 50.0
    m = int(round(spikes * v.size))
    if m:
        pos = rng.choice(v.size, m, replace=False)
        v.reshape(-1)[pos] = rng.choice([-1.0, 1.0], m) * rng.uniform(100, 1000, m)
    return v.astype(dtype)


def weights(rows: int, cols: int, sigma: float, seed: int = 2, dtype=np.float32) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.normal(0.0, sigma, (rows, cols)).astype(dtype)


def rel_to_abs(x: np.ndarray, rel: float) -> float:
    """examples/cuSZp.cpp:176-210 of the reference: float range times the bound."""
    if x.dtype == np.float32:
        return float(np.float32(rel) * (np.float32(x.max()) - np.float32(x.min())))
    return float(np.float32(rel) * np.float32(x.max() - x.min()))


def mixed_field(n: int, seed: int, dtype=np.float32) -> np.ndarray:
    """Small-case generator mixing smooth runs, noise, spikes and exact zeros."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.cumsum(rng.normal(0, 0.3, n)) + rng.normal(0, 0.01, n)
    x[rng.random(n) < 0.02] *= 40
    z = rng.random(n) < 0.01
    x[z] = 0
    if n > 200:
        s = int(rng.integers(0, n - 100))
        x[s:s + 96] = 0.0
    return x.astype(dtype)
