"""Loader for libcuszp_amd.so (the HIP kernels behind the C ABI of include/cuSZp.h).

There is no fallback: if the library is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcuszp_amd.so")

_lib = None


class Uint3(ctypes.Structure):
    """cuszp_dims_t / CUDA uint3: {x, y, z}, x fastest (python/pycuSZp.py:8-9 of the reference)."""

    _fields_ = [("x", ctypes.c_uint), ("y", ctypes.c_uint), ("z", ctypes.c_uint)]


EXPORTED = (
    ["cuSZp_compress", "cuSZp_decompress"]
    + [
        f"cuSZp_{d}_{dim}_{m}_{t}"
        for t in ("f32", "f64")
        for dim in ("1D", "2D", "3D")
        for m in ("fixed", "plain", "outlier")
        for d in ("compress", "decompress")
    ]
    + [
        "cuSZp_max_compressed_size",
        "cuSZp_compress_ex",
        "cuSZp_decompress_ex",
        "cuSZp_compress_async",
        "cuSZp_decompress_async",
        "cuSZp_last_async_status",
        "cuSZp_scan_report",
        "cuSZp_scan_config",
        "cuSZp_debug_fail_allocs",
        "cuSZp_batch_plan_scan_report",
        "cuSZp_status_string",
        "cuSZp_version",
        "cuSZp_batch_plan_create",
        "cuSZp_batch_plan_execute",
        "cuSZp_batch_plan_results",
        "cuSZp_batch_plan_device_sizes",
        "cuSZp_batch_plan_tune",
        "cuSZp_batch_plan_destroy",
        "cuSZp_compress_batch",
        "cuSZp_decompress_batch",
        "cuSZp_server_create",
        "cuSZp_server_submit",
        "cuSZp_server_poll",
        "cuSZp_server_release",
        "cuSZp_server_stats",
        "cuSZp_server_destroy",
        "cuSZp_server_replay",
        "cuSZp_value_range",
        "cuSZp_rel_error_bound",
        "cuSZp_quality",
        "cuSZp_ssim",
        "cuSZp_container_write_header",
        "cuSZp_container_read_header",
        "cuSZp_compress_container",
        "cuSZp_decompress_container",
        "cuSZp_compress_noalloc",
    ]
)

# cuszp_batch_item_t (include/cuSZp.h): 64 bytes
class BatchItem(ctypes.Structure):
    _fields_ = [
        ("d_in", ctypes.c_void_p),
        ("d_out", ctypes.c_void_p),
        ("nbEle", ctypes.c_size_t),
        ("capacity", ctypes.c_size_t),
        ("errorBound", ctypes.c_double),
        ("dim", ctypes.c_int),
        ("dims", Uint3),
        ("type", ctypes.c_int),
        ("mode", ctypes.c_int),
    ]


class ScanReport(ctypes.Structure):
    """cuszp_scan_report_t (include/cuSZp.h, scan diagnostics)."""

    _fields_ = [
        ("info", ctypes.c_uint),
        ("polls", ctypes.c_ulonglong),
        ("waits", ctypes.c_ulonglong),
        ("claims", ctypes.c_ulonglong),
        ("wait_ticks", ctypes.c_ulonglong),
    ]


SCAN_INFO_STOLEN = 8
WARNING_BOUND_EXCEEDED = 8


class ContainerInfo(ctypes.Structure):
    """cuszp_container_info_t (include/cuSZp.h, self-describing container)."""

    _fields_ = [
        ("dim", ctypes.c_int),
        ("type", ctypes.c_int),
        ("mode", ctypes.c_int),
        ("relative", ctypes.c_int),
        ("dims", Uint3),
        ("nbEle", ctypes.c_ulonglong),
        ("errorBound", ctypes.c_double),
        ("relBound", ctypes.c_double),
        ("streamBytes", ctypes.c_ulonglong),
    ]


CONTAINER_HEADER_BYTES = 64


class Completion(ctypes.Structure):
    """cuszp_completion_t (include/cuSZp.h, serving form)."""

    _fields_ = [
        ("id", ctypes.c_ulonglong),
        ("status", ctypes.c_int),
        ("direction", ctypes.c_int),
        ("offset", ctypes.c_size_t),
        ("d_out", ctypes.c_void_p),
        ("bytes", ctypes.c_size_t),
    ]


class Request(ctypes.Structure):
    """cuszp_request_t (include/cuSZp.h, load replay)."""

    _fields_ = [("direction", ctypes.c_int), ("item", BatchItem), ("arrival_s", ctypes.c_double)]


BATCH_COMPRESS, BATCH_DECOMPRESS = 0, 1
BATCH_GRAPH, BATCH_TUNE = 1, 2


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python cuszp_amd/build.py`"
        )
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i32, f32, f64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_float, ctypes.c_double
    psz = ctypes.POINTER(ctypes.c_size_t)
    # generic entries: include/cuSZp.h:38-45 of the reference
    L.cuSZp_compress.argtypes = [vp, vp, sz, psz, f32, i32, Uint3, i32, i32, vp]
    L.cuSZp_compress.restype = None
    L.cuSZp_decompress.argtypes = [vp, vp, sz, sz, f32, i32, Uint3, i32, i32, vp]
    L.cuSZp_decompress.restype = None
    for t, ft in (("f32", f32), ("f64", f64)):
        for m in ("fixed", "plain", "outlier"):
            c = getattr(L, f"cuSZp_compress_1D_{m}_{t}")
            c.argtypes, c.restype = [vp, vp, sz, psz, ft, vp], None
            d = getattr(L, f"cuSZp_decompress_1D_{m}_{t}")
            d.argtypes, d.restype = [vp, vp, sz, sz, ft, vp], None
            for dim in ("2D", "3D"):
                c = getattr(L, f"cuSZp_compress_{dim}_{m}_{t}")
                c.argtypes, c.restype = [vp, vp, sz, psz, Uint3, ft, vp], None
                d = getattr(L, f"cuSZp_decompress_{dim}_{m}_{t}")
                d.argtypes, d.restype = [vp, vp, sz, sz, Uint3, ft, vp], None
    L.cuSZp_max_compressed_size.argtypes = [sz, i32, Uint3, i32, i32]
    L.cuSZp_max_compressed_size.restype = sz
    L.cuSZp_compress_ex.argtypes = [vp, vp, sz, sz, psz, f64, i32, Uint3, i32, i32, vp]
    L.cuSZp_compress_ex.restype = i32
    L.cuSZp_decompress_ex.argtypes = [vp, vp, sz, sz, f64, i32, Uint3, i32, i32, vp]
    L.cuSZp_decompress_ex.restype = i32
    L.cuSZp_compress_async.argtypes = [vp, vp, sz, sz, vp, f64, i32, Uint3, i32, i32, vp]
    L.cuSZp_compress_async.restype = i32
    L.cuSZp_decompress_async.argtypes = [vp, vp, sz, sz, f64, i32, Uint3, i32, i32, vp]
    L.cuSZp_decompress_async.restype = i32
    L.cuSZp_last_async_status.argtypes = [vp]
    L.cuSZp_last_async_status.restype = i32
    L.cuSZp_scan_report.argtypes = [vp, ctypes.POINTER(ScanReport)]
    L.cuSZp_scan_report.restype = i32
    L.cuSZp_scan_config.argtypes = [i32, ctypes.c_uint, ctypes.c_uint, i32]
    L.cuSZp_scan_config.restype = i32
    L.cuSZp_debug_fail_allocs.argtypes = [ctypes.c_uint]
    L.cuSZp_debug_fail_allocs.restype = i32
    L.cuSZp_batch_plan_scan_report.argtypes = [vp, vp, ctypes.POINTER(ScanReport)]
    L.cuSZp_batch_plan_scan_report.restype = i32
    L.cuSZp_status_string.argtypes = [i32]
    L.cuSZp_status_string.restype = ctypes.c_char_p
    L.cuSZp_version.argtypes = []
    L.cuSZp_version.restype = ctypes.c_char_p
    pitems = ctypes.POINTER(BatchItem)
    L.cuSZp_batch_plan_create.argtypes = [ctypes.POINTER(vp), pitems, sz, i32, i32, vp, psz]
    L.cuSZp_batch_plan_create.restype = i32
    L.cuSZp_batch_plan_execute.argtypes = [vp, vp]
    L.cuSZp_batch_plan_execute.restype = i32
    L.cuSZp_batch_plan_results.argtypes = [vp, psz, ctypes.POINTER(i32), vp]
    L.cuSZp_batch_plan_results.restype = i32
    L.cuSZp_batch_plan_device_sizes.argtypes = [vp]
    L.cuSZp_batch_plan_device_sizes.restype = vp
    L.cuSZp_batch_plan_tune.argtypes = [vp, vp, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_uint)]
    L.cuSZp_batch_plan_tune.restype = i32
    L.cuSZp_batch_plan_destroy.argtypes = [vp]
    L.cuSZp_batch_plan_destroy.restype = None
    L.cuSZp_compress_batch.argtypes = [pitems, sz, psz, ctypes.POINTER(i32), vp]
    L.cuSZp_compress_batch.restype = i32
    L.cuSZp_decompress_batch.argtypes = [pitems, sz, ctypes.POINTER(i32), vp]
    L.cuSZp_decompress_batch.restype = i32
    pd = ctypes.POINTER(ctypes.c_double)
    L.cuSZp_server_create.argtypes = [ctypes.POINTER(vp), vp, sz, i32, sz]
    L.cuSZp_server_create.restype = i32
    L.cuSZp_server_submit.argtypes = [vp, i32, pitems, ctypes.POINTER(ctypes.c_ulonglong)]
    L.cuSZp_server_submit.restype = i32
    L.cuSZp_server_poll.argtypes = [vp, ctypes.POINTER(Completion), sz, psz, i32]
    L.cuSZp_server_poll.restype = i32
    L.cuSZp_server_release.argtypes = [vp, ctypes.c_ulonglong]
    L.cuSZp_server_release.restype = i32
    L.cuSZp_server_stats.argtypes = [vp, psz, psz, psz]
    L.cuSZp_server_stats.restype = i32
    L.cuSZp_server_destroy.argtypes = [vp]
    L.cuSZp_server_destroy.restype = None
    L.cuSZp_server_replay.argtypes = [vp, ctypes.POINTER(Request), sz, i32, pd, ctypes.POINTER(i32), pd]
    L.cuSZp_server_replay.restype = i32
    pci = ctypes.POINTER(ContainerInfo)
    L.cuSZp_container_write_header.argtypes = [pci, vp]
    L.cuSZp_container_write_header.restype = i32
    L.cuSZp_container_read_header.argtypes = [vp, sz, pci]
    L.cuSZp_container_read_header.restype = i32
    L.cuSZp_compress_container.argtypes = [vp, vp, sz, sz, psz, f64, f64, i32, Uint3, i32, i32, vp]
    L.cuSZp_compress_container.restype = i32
    L.cuSZp_decompress_container.argtypes = [vp, sz, vp, sz, pci, vp]
    L.cuSZp_decompress_container.restype = i32
    L.cuSZp_compress_noalloc.argtypes = [vp, vp, sz, sz, vp, f64, i32, Uint3, i32, i32, vp]
    L.cuSZp_compress_noalloc.restype = i32
    L.cuSZp_value_range.argtypes = [vp, sz, i32, ctypes.POINTER(f64), ctypes.POINTER(f64), vp]
    L.cuSZp_value_range.restype = i32
    L.cuSZp_rel_error_bound.argtypes = [vp, sz, i32, f32, ctypes.POINTER(f32), vp]
    L.cuSZp_rel_error_bound.restype = i32
    L.cuSZp_quality.argtypes = [vp, vp, sz, i32, ctypes.POINTER(f64), vp]
    L.cuSZp_quality.restype = i32
    L.cuSZp_ssim.argtypes = [vp, vp, Uint3, i32, ctypes.POINTER(f64), vp]
    L.cuSZp_ssim.restype = i32
    _lib = L
    return L
