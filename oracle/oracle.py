"""ctypes binding of the C++ cuSZp oracle (oracle/cuszp_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product package (cuszp_amd)
never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_cuszp.so")

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (seconds)."""
    if force or not os.path.exists(LIB_PATH) or (
        os.path.getmtime(LIB_PATH) < os.path.getmtime(os.path.join(HERE, "cuszp_oracle.cpp"))
    ):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u32, sz, dbl, i32, vp = ctypes.c_uint32, ctypes.c_size_t, ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.cuszp_oracle_quantize_f32.argtypes = [ctypes.c_float, ctypes.c_float]
        L.cuszp_oracle_quantize_f32.restype = ctypes.c_int32
        L.cuszp_oracle_quantize_f64.argtypes = [dbl, dbl]
        L.cuszp_oracle_quantize_f64.restype = ctypes.c_int32
        L.cuszp_oracle_recip_f32.argtypes = [ctypes.c_float]
        L.cuszp_oracle_recip_f32.restype = ctypes.c_float
        L.cuszp_oracle_recip_f64.argtypes = [dbl]
        L.cuszp_oracle_recip_f64.restype = dbl
        for name in ("cuszp_oracle_header_size", "cuszp_oracle_num_blocks", "cuszp_oracle_max_size"):
            f = getattr(L, name)
            f.argtypes = [i32, sz, u32, u32, u32]
            f.restype = sz
        L.cuszp_oracle_compress.argtypes = [vp, i32, i32, i32, sz, u32, u32, u32, dbl, vp, sz, i32]
        L.cuszp_oracle_compress.restype = ctypes.c_longlong
        L.cuszp_oracle_decompress.argtypes = [vp, i32, i32, i32, sz, u32, u32, u32, dbl, vp, sz, i32]
        L.cuszp_oracle_decompress.restype = i32
        L.cuszp_oracle_rel_eb.argtypes = [vp, i32, sz, ctypes.c_float, ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
        L.cuszp_oracle_rel_eb.restype = ctypes.c_float
        L.cuszp_oracle_quality.argtypes = [vp, vp, i32, sz, ctypes.POINTER(dbl)]
        L.cuszp_oracle_quality.restype = None
        L.cuszp_oracle_ssim.argtypes = [vp, vp, i32, sz, sz, sz]
        L.cuszp_oracle_ssim.restype = dbl
        L.cuszp_oracle_max_threads.argtypes = []
        L.cuszp_oracle_max_threads.restype = i32
        _lib = L
    return _lib


def _dims(dim, n, dims):
    if dim == 1:
        return (0, 0, 0)
    return tuple(int(v) for v in dims)


def quantize_f32(x: float, r: float) -> int:
    return lib().cuszp_oracle_quantize_f32(x, r)


def quantize_f64(x: float, r: float) -> int:
    return lib().cuszp_oracle_quantize_f64(x, r)


def header_size(dim, n, dims=(0, 0, 0)):
    return lib().cuszp_oracle_header_size(dim, n, *_dims(dim, n, dims))


def num_blocks(dim, n, dims=(0, 0, 0)):
    return lib().cuszp_oracle_num_blocks(dim, n, *_dims(dim, n, dims))


def max_size(dim, n, dims=(0, 0, 0)):
    return lib().cuszp_oracle_max_size(dim, n, *_dims(dim, n, dims))


def rel_eb(data: np.ndarray, rel: float):
    """(abs eb, min, max) as the reference CLI derives a relative bound (examples/cuSZp.cpp:176-210)."""
    data = np.ascontiguousarray(data)
    typ = 0 if data.dtype == np.float32 else 1
    mn, mx = ctypes.c_double(0), ctypes.c_double(0)
    eb = lib().cuszp_oracle_rel_eb(data.ctypes.data, typ, data.size, rel, ctypes.byref(mn), ctypes.byref(mx))
    return float(eb), mn.value, mx.value


QUALITY_FIELDS = ("mse", "range", "psnr", "norm_err", "norm_err_norm", "nrmse", "max_abs_err", "max_rel_err",
                  "correlation")


def quality(ori: np.ndarray, dec: np.ndarray) -> dict:
    """Reconstruction quality as src/cuSZp_utility.cu:891-950 computes it (sequential)."""
    ori = np.ascontiguousarray(ori)
    dec = np.ascontiguousarray(dec, dtype=ori.dtype)
    out = (ctypes.c_double * 9)()
    lib().cuszp_oracle_quality(ori.ctypes.data, dec.ctypes.data, 0 if ori.dtype == np.float32 else 1, ori.size, out)
    return dict(zip(QUALITY_FIELDS, list(out)))


def ssim(ori: np.ndarray, dec: np.ndarray, dims) -> float:
    """Mean SSIM as src/cuSZp_utility.cu:761-889 computes it; dims = (x, y, z), x fastest."""
    ori = np.ascontiguousarray(ori)
    dec = np.ascontiguousarray(dec, dtype=ori.dtype)
    x, y, z = dims
    return lib().cuszp_oracle_ssim(ori.ctypes.data, dec.ctypes.data, 0 if ori.dtype == np.float32 else 1, x, y, z)


def max_threads():
    return lib().cuszp_oracle_max_threads()


def compress(data: np.ndarray, eb: float, mode: int, dim: int = 1, dims=(0, 0, 0), nthreads: int = 0) -> bytes:
    """Compress a host array. dims = (x, y, z) with x fastest, as in cuSZp's uint3."""
    data = np.ascontiguousarray(data)
    if data.dtype == np.float32:
        typ = 0
        eb = float(np.float32(eb))
    elif data.dtype == np.float64:
        typ = 1
    else:
        raise TypeError(data.dtype)
    n = data.size
    if nthreads <= 0:
        nthreads = max_threads()
    cap = max_size(dim, n, dims)
    out = np.empty(max(cap, 1), dtype=np.uint8)
    r = lib().cuszp_oracle_compress(
        data.ctypes.data, typ, dim, mode, n, *_dims(dim, n, dims), eb, out.ctypes.data, cap, nthreads
    )
    if r < 0:
        raise ValueError("oracle compress rejected the arguments")
    return out[:r].tobytes()


def decompress(cmp: bytes, n: int, dtype, eb: float, mode: int, dim: int = 1, dims=(0, 0, 0), nthreads: int = 0) -> np.ndarray:
    dtype = np.dtype(dtype)
    typ = 0 if dtype == np.float32 else 1
    if typ == 0:
        eb = float(np.float32(eb))
    if nthreads <= 0:
        nthreads = max_threads()
    out = np.empty(n, dtype=dtype)
    buf = np.frombuffer(cmp, dtype=np.uint8) if len(cmp) else np.zeros(1, np.uint8)
    rc = lib().cuszp_oracle_decompress(
        out.ctypes.data, typ, dim, mode, n, *_dims(dim, n, dims), eb, buf.ctypes.data, len(cmp), nthreads
    )
    if rc != 0:
        raise ValueError(f"oracle decompress failed ({rc})")
    return out
