"""Shared helpers: call the HIP library through its C ABI on host arrays, and
the error-bound tolerance used by the parity tests."""
from __future__ import annotations

import numpy as np

TYPE_OF = {np.dtype(np.float32): 0, np.dtype(np.float64): 1}
UNIT_ROUNDOFF = {np.dtype(np.float32): 2.0**-24, np.dtype(np.float64): 2.0**-53}


def gpu_compress(x: np.ndarray, eb: float, mode: int, dim: int = 1, dims=(0, 0, 0), stream=None) -> bytes:
    import torch
    from cuszp_amd import pycuSZp as P

    t = torch.from_numpy(np.ascontiguousarray(x).ravel()).cuda()
    typ = TYPE_OF[x.dtype]
    cap = P.max_compressed_size(t.numel(), dim, dims, typ, mode)
    out = torch.full((max(cap, 1),), 0xA5, dtype=torch.uint8, device=t.device)
    n = P.compress_ex(t, out, cap, t.numel(), eb, dim, dims, typ, mode, stream)
    torch.cuda.synchronize()
    return out[:n].cpu().numpy().tobytes()


def gpu_decompress(cmp: bytes, n: int, dtype, eb: float, mode: int, dim: int = 1, dims=(0, 0, 0),
                   fill=np.nan) -> np.ndarray:
    import torch
    from cuszp_amd import pycuSZp as P

    dtype = np.dtype(dtype)
    c = torch.from_numpy(np.frombuffer(cmp, np.uint8).copy() if len(cmp) else np.zeros(1, np.uint8)).cuda()
    out = torch.full((max(n, 1),), float(fill), dtype=torch.float32 if dtype == np.float32 else torch.float64,
                     device=c.device)
    P.decompress_ex(out, c, n, len(cmp), eb, dim, dims, TYPE_OF[dtype], mode)
    torch.cuda.synchronize()
    return out[:n].cpu().numpy()


def error_tolerance(x: np.ndarray, xhat: np.ndarray, eb: float) -> np.ndarray:
    """Per-element bound on |x - xhat| for the reference's arithmetic.

    The decoder returns RN(RN(q*eb)*2) for q = round(RN(x*r)), r = RN(0.5/eb),
    so |x - xhat| <= eb + (|x| + |xhat|) * 2u + eb*2^-20 where u is the unit
    roundoff of the element type (three relative roundings, each at most u).
    Elements whose quantisation code saturates (|x|*r >= 2^31) are excluded by
    the callers.
    """
    u = UNIT_ROUNDOFF[x.dtype]
    ax = np.abs(x.astype(np.float64))
    axh = np.abs(xhat.astype(np.float64))
    return eb * (1 + 2.0**-20) + (ax + axh) * 2 * u


def check_error_bound(x: np.ndarray, xhat: np.ndarray, eb: float):
    x64 = x.astype(np.float64)
    finite = np.isfinite(x64) & (np.abs(x64) * (0.5 / eb) < 2.0**30)
    err = np.abs(x64 - xhat.astype(np.float64))
    tol = error_tolerance(x, xhat, eb)
    bad = finite & ~(err <= tol)
    assert not bad.any(), f"{bad.sum()} elements beyond tolerance, worst err/eb={np.max(err[finite] / eb):.6f}"
    return float(np.max(err[finite] / eb)) if finite.any() else 0.0
