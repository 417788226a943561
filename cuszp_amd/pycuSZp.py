"""cuSZp's Python binding (python/pycuSZp.py of the reference), re-pointed at
libcuszp_amd.so and PyTorch-ROCm tensors.

`cuSZp.compress` / `cuSZp.decompress` keep the reference's argument names,
order, defaults and return values (python/pycuSZp.py:38-77). Pointers may be
given as integers (e.g. `tensor.data_ptr()`), ctypes pointers, or torch tensors;
`stream` may be None (null stream), an integer handle or a torch.cuda.Stream.

On top of that, `compress_tensor` / `decompress_tensor` allocate outputs on the
tensor's device and return them, and `rel_to_abs` turns a relative bound into
the absolute one exactly as the reference CLI does (examples/cuSZp.cpp:176-210).
"""
from __future__ import annotations

import ctypes
import warnings
from typing import Optional, Tuple

from ._lib import SCAN_INFO_STOLEN, WARNING_BOUND_EXCEEDED, ScanReport, Uint3, load

lib = load()

CUSZP_DIM_1D, CUSZP_DIM_2D, CUSZP_DIM_3D = 1, 2, 3
CUSZP_MODE_FIXED, CUSZP_MODE_PLAIN, CUSZP_MODE_OUTLIER = 0, 1, 2
CUSZP_TYPE_FLOAT, CUSZP_TYPE_DOUBLE = 0, 1
MODES = {"fixed": 0, "plain": 1, "outlier": 2}


class BoundExceededWarning(UserWarning):
    """The stream is complete, but a 2D/3D outlier-mode block holds a residual
    of magnitude 2^31 and decodes to zeros (status CUSZP_WARNING_BOUND_EXCEEDED)."""


class CuSZpError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = lib.cuSZp_status_string(status).decode()
        super().__init__(f"{what}: {msg}" if what else msg)


def _ptr(p):
    if p is None:
        return None
    if hasattr(p, "data_ptr"):
        return ctypes.c_void_p(p.data_ptr())
    if isinstance(p, ctypes.c_void_p):
        return p
    return ctypes.c_void_p(int(p))


def _stream(s):
    if s is None:
        return ctypes.c_void_p(0)
    if hasattr(s, "cuda_stream"):
        return ctypes.c_void_p(s.cuda_stream)
    if isinstance(s, ctypes.c_void_p):
        return s
    return ctypes.c_void_p(int(s))


class cuSZp:
    """Same surface as the reference's `cuSZp` class (python/pycuSZp.py:38)."""

    def __init__(self):
        pass

    def compress(self, d_oriData, d_cmpBytes, num_elements, error_bound=0.01,
                 dim=1, dims=(0, 0, 0), data_type=0, mode=0, stream=None):
        compressed_size = ctypes.c_size_t(0)
        dims_struct = Uint3(dims[0], dims[1], dims[2])
        lib.cuSZp_compress(
            _ptr(d_oriData),
            _ptr(d_cmpBytes),
            num_elements,
            ctypes.byref(compressed_size),
            ctypes.c_float(error_bound),
            ctypes.c_int(dim),
            dims_struct,
            ctypes.c_int(data_type),
            ctypes.c_int(mode),
            _stream(stream),
        )
        return compressed_size.value

    def decompress(self, d_decData, d_cmpBytes, num_elements, compressed_size,
                   error_bound=0.01, dim=1, dims=(0, 0, 0), data_type=0, mode=0, stream=None):
        dims_struct = Uint3(dims[0], dims[1], dims[2])
        lib.cuSZp_decompress(
            _ptr(d_decData),
            _ptr(d_cmpBytes),
            num_elements,
            compressed_size,
            ctypes.c_float(error_bound),
            ctypes.c_int(dim),
            dims_struct,
            ctypes.c_int(data_type),
            ctypes.c_int(mode),
            _stream(stream),
        )


# ---------------------------------------------------------------------------
# status-checked helpers
# ---------------------------------------------------------------------------
def max_compressed_size(num_elements: int, dim: int = 1, dims=(0, 0, 0), data_type: int = 0, mode: int = 0) -> int:
    return lib.cuSZp_max_compressed_size(num_elements, dim, Uint3(*dims), data_type, mode)


def compress_ex(d_oriData, d_cmpBytes, capacity: int, num_elements: int, error_bound: float,
                dim: int = 1, dims=(0, 0, 0), data_type: int = 0, mode: int = 0, stream=None) -> int:
    size = ctypes.c_size_t(0)
    rc = lib.cuSZp_compress_ex(_ptr(d_oriData), _ptr(d_cmpBytes), capacity, num_elements, ctypes.byref(size),
                               float(error_bound), dim, Uint3(*dims), data_type, mode, _stream(stream))
    if rc == WARNING_BOUND_EXCEEDED:
        warnings.warn(lib.cuSZp_status_string(rc).decode(), BoundExceededWarning, stacklevel=2)
    elif rc:
        raise CuSZpError(rc, "compress")
    return size.value


def scan_report(stream=None) -> dict:
    """Scan diagnostics of the single-array entries on `stream` since the last
    report (cuSZp_scan_report): info bits and, while the stats switch is on,
    lookback polls, waits, claims and wait time in microseconds."""
    r = ScanReport()
    rc = lib.cuSZp_scan_report(_stream(stream), ctypes.byref(r))
    if rc:
        raise CuSZpError(rc, "scan report")
    return _report_dict(r)


def _report_dict(r) -> dict:
    return {"stolen": bool(r.info & SCAN_INFO_STOLEN), "info": r.info, "polls": r.polls, "waits": r.waits,
            "claims": r.claims, "wait_us": r.wait_ticks / 100.0}


def scan_config(steal_ticks: int = -1, hold_every: int = 0, hold_ticks: int = 0, stats: bool = False) -> None:
    """Scan settings of later launches (cuSZp_scan_config); the defaults restore
    normal operation."""
    rc = lib.cuSZp_scan_config(int(steal_ticks), int(hold_every), int(hold_ticks), 1 if stats else 0)
    if rc:
        raise CuSZpError(rc, "scan config")


def decompress_ex(d_decData, d_cmpBytes, num_elements: int, compressed_size: int, error_bound: float,
                  dim: int = 1, dims=(0, 0, 0), data_type: int = 0, mode: int = 0, stream=None) -> None:
    rc = lib.cuSZp_decompress_ex(_ptr(d_decData), _ptr(d_cmpBytes), num_elements, compressed_size,
                                 float(error_bound), dim, Uint3(*dims), data_type, mode, _stream(stream))
    if rc:
        raise CuSZpError(rc, "decompress")


def _type_of(t) -> int:
    import torch

    if t.dtype == torch.float32:
        return CUSZP_TYPE_FLOAT
    if t.dtype == torch.float64:
        return CUSZP_TYPE_DOUBLE
    raise TypeError(f"cuSZp compresses float32/float64 tensors, got {t.dtype}")


def value_range(t, stream=None) -> Tuple[float, float]:
    """(min, max) of a contiguous float32/float64 CUDA tensor, computed on the
    device as the reference CLI's loop does (examples/cuSZp.cpp:176-210): NaNs
    are skipped unless element 0 is NaN. Synchronises the stream."""
    if not t.is_contiguous() or t.numel() == 0:
        raise ValueError("tensor must be contiguous and non-empty")
    mn, mx = ctypes.c_double(0), ctypes.c_double(0)
    rc = lib.cuSZp_value_range(_ptr(t), t.numel(), _type_of(t), ctypes.byref(mn), ctypes.byref(mx), _stream(stream))
    if rc:
        raise CuSZpError(rc, "value range")
    return mn.value, mx.value


def rel_to_abs(t, rel_eb: float, stream=None) -> float:
    """Relative -> absolute error bound like examples/cuSZp.cpp:176-210, on the
    device: f32: (float)rel * (float)(max - min) in float; f64: rel * (float)(max - min)."""
    if not t.is_contiguous() or t.numel() == 0:
        raise ValueError("tensor must be contiguous and non-empty")
    eb = ctypes.c_float(0)
    rc = lib.cuSZp_rel_error_bound(_ptr(t), t.numel(), _type_of(t), float(rel_eb), ctypes.byref(eb), _stream(stream))
    if rc:
        raise CuSZpError(rc, "relative error bound")
    return eb.value


QUALITY_FIELDS = ("mse", "range", "psnr", "norm_err", "norm_err_norm", "nrmse", "max_abs_err", "max_rel_err",
                  "correlation")


def quality(ori, dec, stream=None) -> dict:
    """Reconstruction quality of `dec` against `ori` (same dtype and size,
    contiguous CUDA tensors) as src/cuSZp_utility.cu:891-950 computes it, on
    the device. Synchronises the stream."""
    if ori.dtype != dec.dtype or ori.numel() != dec.numel() or not (ori.is_contiguous() and dec.is_contiguous()):
        raise ValueError("ori and dec must be contiguous tensors of one dtype and size")
    if ori.numel() == 0:
        raise ValueError("empty tensors")
    out = (ctypes.c_double * 9)()
    rc = lib.cuSZp_quality(_ptr(ori), _ptr(dec), ori.numel(), _type_of(ori), out, _stream(stream))
    if rc:
        raise CuSZpError(rc, "quality")
    return dict(zip(QUALITY_FIELDS, list(out)))


def ssim(ori, dec, dims, stream=None) -> float:
    """Mean SSIM over 7x7x7 windows moved by 2 (src/cuSZp_utility.cu:761-889),
    on the device; dims = (x, y, z), x fastest, each at least 7."""
    if ori.dtype != dec.dtype or ori.numel() != dec.numel() or not (ori.is_contiguous() and dec.is_contiguous()):
        raise ValueError("ori and dec must be contiguous tensors of one dtype and size")
    if ori.numel() != dims[0] * dims[1] * dims[2]:
        raise ValueError("dims do not match the number of values")
    out = ctypes.c_double(0)
    rc = lib.cuSZp_ssim(_ptr(ori), _ptr(dec), Uint3(*dims), _type_of(ori), ctypes.byref(out), _stream(stream))
    if rc:
        raise CuSZpError(rc, "ssim")
    return out.value


def compress_tensor_rel(t, rel_eb: float, mode="plain", dim: int = 1, dims: Optional[Tuple[int, int, int]] = None,
                        stream=None):
    """compress_tensor with a relative bound: returns (uint8 tensor, size, absolute eb)."""
    eb = rel_to_abs(t, rel_eb, stream)
    out, n = compress_tensor(t, eb, mode, dim, dims, stream)
    return out, n, eb


def compress_tensor(t, error_bound: float, mode="plain", dim: int = 1, dims: Optional[Tuple[int, int, int]] = None,
                    stream=None):
    """Compress a contiguous float32/float64 CUDA tensor. Returns (uint8 tensor, size)."""
    import torch

    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    m = MODES[mode] if isinstance(mode, str) else int(mode)
    typ = _type_of(t)
    dims = tuple(dims) if dims is not None else (0, 0, 0)
    cap = max_compressed_size(t.numel(), dim, dims, typ, m)
    out = torch.empty(max(cap, 1), dtype=torch.uint8, device=t.device)
    if stream is None:
        stream = torch.cuda.current_stream(t.device)
    n = compress_ex(t, out, cap, t.numel(), error_bound, dim, dims, typ, m, stream)
    return out, n


def decompress_tensor(cmp, size: int, num_elements: int, dtype, error_bound: float, mode="plain", dim: int = 1,
                      dims: Optional[Tuple[int, int, int]] = None, stream=None):
    import torch

    m = MODES[mode] if isinstance(mode, str) else int(mode)
    out = torch.empty(num_elements, dtype=dtype, device=cmp.device)
    typ = _type_of(out)
    dims = tuple(dims) if dims is not None else (0, 0, 0)
    if stream is None:
        stream = torch.cuda.current_stream(cmp.device)
    decompress_ex(out, cmp, num_elements, size, error_bound, dim, dims, typ, m, stream)
    return out


# ---------------------------------------------------------------------------
# self-describing container: a 64-byte header (type, mode, dims, bound,
# length; layout in include/cuSZp.h) in front of the unchanged raw stream
# ---------------------------------------------------------------------------
from ._lib import CONTAINER_HEADER_BYTES, ContainerInfo  # noqa: E402


def _info_dict(ci) -> dict:
    return {"dim": ci.dim, "type": ci.type, "mode": ci.mode, "relative": bool(ci.relative),
            "dims": (ci.dims.x, ci.dims.y, ci.dims.z), "num_elements": ci.nbEle, "error_bound": ci.errorBound,
            "rel_bound": ci.relBound, "stream_bytes": ci.streamBytes}


def read_container_header(data: bytes) -> dict:
    """Decodes and validates the header at the start of `data` (host bytes)."""
    ci = ContainerInfo()
    buf = (ctypes.c_ubyte * max(len(data), 1)).from_buffer_copy(data if data else b"\0")
    rc = lib.cuSZp_container_read_header(buf, len(data), ctypes.byref(ci))
    if rc:
        raise CuSZpError(rc, "container header")
    return _info_dict(ci)


def compress_container(t, error_bound: float, mode="plain", dim: int = 1,
                       dims: Optional[Tuple[int, int, int]] = None, rel_bound: float = 0.0, stream=None):
    """Compress a contiguous float32/float64 CUDA tensor into a container.
    Returns (uint8 tensor, container bytes); the raw stream starts at byte 64.
    With rel_bound > 0 the header records it (error_bound is then the absolute
    bound derived from it, e.g. rel_to_abs(t, rel_bound))."""
    import torch

    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    m = MODES[mode] if isinstance(mode, str) else int(mode)
    typ = _type_of(t)
    dims = tuple(dims) if dims is not None else (0, 0, 0)
    cap = CONTAINER_HEADER_BYTES + max_compressed_size(t.numel(), dim, dims, typ, m)
    out = torch.empty(cap, dtype=torch.uint8, device=t.device)
    if stream is None:
        stream = torch.cuda.current_stream(t.device)
    size = ctypes.c_size_t(0)
    rc = lib.cuSZp_compress_container(_ptr(t), _ptr(out), cap, t.numel(), ctypes.byref(size), float(error_bound),
                                      float(rel_bound), dim, Uint3(*dims), typ, m, _stream(stream))
    if rc == WARNING_BOUND_EXCEEDED:
        warnings.warn(lib.cuSZp_status_string(rc).decode(), BoundExceededWarning, stacklevel=2)
    elif rc:
        raise CuSZpError(rc, "compress container")
    return out, size.value


def decompress_container(buf, size: Optional[int] = None, stream=None):
    """Decompress a container held in a CUDA uint8 tensor (the first `size`
    bytes; all of it by default). Returns (values tensor, header dict); the
    values' dtype, count and shape come from the header alone."""
    import torch

    size = buf.numel() if size is None else int(size)
    if size < CONTAINER_HEADER_BYTES:
        raise CuSZpError(5, "decompress container")
    info = read_container_header(buf[:CONTAINER_HEADER_BYTES].cpu().numpy().tobytes())
    dtype = torch.float64 if info["type"] == CUSZP_TYPE_DOUBLE else torch.float32
    out = torch.empty(max(info["num_elements"], 1), dtype=dtype, device=buf.device)
    if stream is None:
        stream = torch.cuda.current_stream(buf.device)
    ci = ContainerInfo()
    rc = lib.cuSZp_decompress_container(_ptr(out), out.numel(), _ptr(buf), size, ctypes.byref(ci), _stream(stream))
    if rc:
        raise CuSZpError(rc, "decompress container")
    return out[:info["num_elements"]], _info_dict(ci)


# ---------------------------------------------------------------------------
# batched form: many arrays in one launch per direction (include/cuSZp.h,
# cuSZp_batch_plan_*). Every array's stream is byte-identical to its
# single-array stream.
# ---------------------------------------------------------------------------
from ._lib import BATCH_COMPRESS, BATCH_DECOMPRESS, BATCH_GRAPH, BATCH_TUNE, BatchItem  # noqa: E402


def batch_item(d_in, d_out, num_elements: int, capacity: int, error_bound: float, dim: int = 1,
               dims=(0, 0, 0), data_type: int = 0, mode: int = 0) -> BatchItem:
    it = BatchItem()
    it.d_in = _ptr(d_in).value if d_in is not None else None
    it.d_out = _ptr(d_out).value if d_out is not None else None
    it.nbEle = num_elements
    it.capacity = capacity
    it.errorBound = float(error_bound)
    it.dim = dim
    it.dims = Uint3(*dims)
    it.type = data_type
    it.mode = mode
    return it


class BatchPlan:
    """Owns a cuSZp_batch_plan_t. `execute()` enqueues the whole batch (a graph
    replay when created with graph=True); `results()` synchronises and returns
    per-array stream lengths (compress) and statuses."""

    def __init__(self, items, direction: int = BATCH_COMPRESS, graph: bool = False, tune: bool = False,
                 stream=None):
        self.count = len(items)
        self._items = (BatchItem * max(self.count, 1))(*items)
        self._plan = ctypes.c_void_p(0)
        bad = ctypes.c_size_t(0)
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_TUNE if tune else 0)
        rc = lib.cuSZp_batch_plan_create(ctypes.byref(self._plan), self._items, self.count, direction, flags,
                                         _stream(stream), ctypes.byref(bad))
        if rc:
            raise CuSZpError(rc, f"batch plan (array {bad.value})")
        self.direction = direction

    def execute(self, stream=None) -> None:
        rc = lib.cuSZp_batch_plan_execute(self._plan, _stream(stream))
        if rc:
            raise CuSZpError(rc, "batch execute")

    def results(self, stream=None, check: bool = True):
        sizes = (ctypes.c_size_t * max(self.count, 1))()
        st = (ctypes.c_int * max(self.count, 1))()
        rc = lib.cuSZp_batch_plan_results(self._plan, sizes, st, _stream(stream))
        if rc == WARNING_BOUND_EXCEEDED:
            warnings.warn(lib.cuSZp_status_string(rc).decode(), BoundExceededWarning, stacklevel=2)
        elif rc and check:
            raise CuSZpError(rc, "batch results")
        return list(sizes)[: self.count], list(st)[: self.count]

    def tune(self, stream=None):
        w = ctypes.c_int(0)
        g = ctypes.c_uint(0)
        rc = lib.cuSZp_batch_plan_tune(self._plan, _stream(stream), ctypes.byref(w), ctypes.byref(g))
        if rc:
            raise CuSZpError(rc, "batch tune")
        return w.value, g.value

    def device_sizes(self) -> int:
        return lib.cuSZp_batch_plan_device_sizes(self._plan) or 0

    def scan_report(self, stream=None) -> dict:
        r = ScanReport()
        rc = lib.cuSZp_batch_plan_scan_report(self._plan, _stream(stream), ctypes.byref(r))
        if rc:
            raise CuSZpError(rc, "batch scan report")
        return _report_dict(r)

    def close(self) -> None:
        if self._plan:
            lib.cuSZp_batch_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p(0)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compress_tensors(tensors, error_bounds, modes, dims_list=None, dims_kind=None, stream=None, graph=False,
                     tune=False):
    """Compress a list of contiguous float32/float64 CUDA tensors in one batched
    launch. error_bounds / modes: one per tensor (absolute bounds; mode names or
    ints). dims_kind[i] in {1, 2, 3} with dims_list[i] = (x, y, z) for 2D/3D.
    Returns [(uint8 tensor, size), ...]."""
    import torch

    n = len(tensors)
    dims_kind = dims_kind or [1] * n
    dims_list = dims_list or [(0, 0, 0)] * n
    outs, items = [], []
    for t, eb, m, dk, dd in zip(tensors, error_bounds, modes, dims_kind, dims_list):
        if not t.is_contiguous():
            raise ValueError("tensors must be contiguous")
        mi = MODES[m] if isinstance(m, str) else int(m)
        typ = _type_of(t)
        cap = max_compressed_size(t.numel(), dk, tuple(dd), typ, mi)
        out = torch.empty(max(cap, 1), dtype=torch.uint8, device=t.device)
        outs.append(out)
        items.append(batch_item(t, out, t.numel(), cap, eb, dk, tuple(dd), typ, mi))
    if stream is None and n:
        stream = torch.cuda.current_stream(tensors[0].device)
    with BatchPlan(items, BATCH_COMPRESS, graph=graph, tune=tune, stream=stream) as plan:
        plan.execute(stream)
        sizes, _ = plan.results(stream)
    return list(zip(outs, sizes))


# ---------------------------------------------------------------------------
# serving form: a request server with continuous batching (include/cuSZp.h,
# cuSZp_server_*). Outputs are allocated in a caller-provided device pool.
# ---------------------------------------------------------------------------
from ._lib import Completion, Request  # noqa: E402


class Server:
    """Owns a cuszp_server_t over `pool` (a contiguous CUDA uint8 tensor, kept
    alive by the server). `submit_*` return a request id; `poll()` returns
    completed requests as (id, status, direction, offset, nbytes) tuples, whose
    output is pool[offset:offset + nbytes] until `release(id)`."""

    def __init__(self, pool, lanes: int = 3, max_batch: int = 256):
        if not pool.is_contiguous():
            raise ValueError("pool must be contiguous")
        self.pool = pool
        self._s = ctypes.c_void_p(0)
        rc = lib.cuSZp_server_create(ctypes.byref(self._s), pool.data_ptr(), pool.numel() * pool.element_size(),
                                     lanes, max_batch)
        if rc:
            raise CuSZpError(rc, "server create")
        self._buf = (Completion * 256)()

    def submit(self, direction: int, item) -> int:
        rid = ctypes.c_ulonglong(0)
        rc = lib.cuSZp_server_submit(self._s, direction, ctypes.byref(item), ctypes.byref(rid))
        if rc:
            raise CuSZpError(rc, "server submit")
        return rid.value

    def submit_compress(self, t, error_bound: float, mode=1, dim: int = 1, dims=(0, 0, 0)) -> int:
        mi = MODES[mode] if isinstance(mode, str) else int(mode)
        return self.submit(BATCH_COMPRESS, batch_item(t if t.numel() else None, None, t.numel(), 0, error_bound,
                                                      dim, tuple(dims), _type_of(t), mi))

    def submit_decompress(self, cmp, cmp_size: int, num_elements: int, dtype, error_bound: float, mode=1,
                          dim: int = 1, dims=(0, 0, 0)) -> int:
        import torch

        mi = MODES[mode] if isinstance(mode, str) else int(mode)
        typ = CUSZP_TYPE_FLOAT if dtype == torch.float32 else CUSZP_TYPE_DOUBLE
        return self.submit(BATCH_DECOMPRESS, batch_item(cmp if num_elements else None, None, num_elements, cmp_size,
                                                        error_bound, dim, tuple(dims), typ, mi))

    def poll(self, wait: bool = False):
        n = ctypes.c_size_t(0)
        out = []
        while True:
            rc = lib.cuSZp_server_poll(self._s, self._buf, len(self._buf), ctypes.byref(n), 1 if wait and not out else 0)
            if rc:
                raise CuSZpError(rc, "server poll")
            out += [(c.id, c.status, c.direction, c.offset, c.bytes) for c in self._buf[: n.value]]
            if n.value < len(self._buf):
                return out

    def release(self, rid: int) -> None:
        rc = lib.cuSZp_server_release(self._s, rid)
        if rc:
            raise CuSZpError(rc, "server release")

    def stats(self):
        q, f, u = ctypes.c_size_t(0), ctypes.c_size_t(0), ctypes.c_size_t(0)
        rc = lib.cuSZp_server_stats(self._s, ctypes.byref(q), ctypes.byref(f), ctypes.byref(u))
        if rc:
            raise CuSZpError(rc, "server stats")
        return q.value, f.value, u.value

    def replay(self, requests, sequential: bool = False):
        """requests: [(direction, BatchItem, arrival_s)] with non-decreasing
        arrivals. Returns (latencies_s, statuses, makespan_s)."""
        n = len(requests)
        arr = (Request * max(n, 1))()
        for i, (d, it, t) in enumerate(requests):
            arr[i].direction = d
            arr[i].item = it
            arr[i].arrival_s = t
        lat = (ctypes.c_double * max(n, 1))()
        st = (ctypes.c_int * max(n, 1))()
        mk = ctypes.c_double(0)
        rc = lib.cuSZp_server_replay(self._s, arr, n, 1 if sequential else 0, lat, st, ctypes.byref(mk))
        if rc:
            raise CuSZpError(rc, "server replay")
        return list(lat)[:n], list(st)[:n], mk.value

    def close(self) -> None:
        if self._s:
            lib.cuSZp_server_destroy(self._s)
            self._s = ctypes.c_void_p(0)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
