"""cuSZp_compress_noalloc (the entry that allocates nothing: its scan state is
a fixed device array of the library) against cuSZp_compress_ex, the pooled
entry it mirrors: identical streams and lengths, a success status in the
result word, on two calls in a row with the second input smaller (the second
call must not see the first call's scan state), for each dimensionality and
both types; plus the size limit and the empty input.
"""
import ctypes

import numpy as np
import pytest

from tests.workloads import grf_volume, mixed_field

pytestmark = pytest.mark.gpu

TOO_LARGE = 3


def _load():
    from cuszp_amd import _lib

    L = _lib.load()
    vp, sz, i32, f64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
    L.cuSZp_compress_noalloc.argtypes = [vp, vp, sz, sz, vp, f64, i32, _lib.Uint3, i32, i32, vp]
    L.cuSZp_compress_noalloc.restype = i32
    return L, _lib


def _pair(L, _lib, torch, device, x, eb, dim, dims, typ, mode):
    """(noalloc stream, noalloc result words, pooled stream) for host array x."""
    n = x.size
    U = _lib.Uint3(*dims)
    cap = L.cuSZp_max_compressed_size(n, dim, U, typ, mode)
    xd = torch.from_numpy(x).to(device)
    out_a = torch.full((cap,), 0xA5, dtype=torch.uint8, device=device)
    out_b = torch.full((cap,), 0x5A, dtype=torch.uint8, device=device)
    res = torch.full((2,), -1, dtype=torch.int64, device=device)
    stream = torch.cuda.current_stream(device)
    sp = ctypes.c_void_p(stream.cuda_stream)
    rc = L.cuSZp_compress_noalloc(ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(out_a.data_ptr()), cap, n,
                                  ctypes.c_void_p(res.data_ptr()), eb, dim, U, typ, mode, sp)
    assert rc == 0, L.cuSZp_status_string(rc)
    torch.cuda.synchronize()
    size = ctypes.c_size_t(0)
    rc = L.cuSZp_compress_ex(ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(out_b.data_ptr()), cap, n,
                             ctypes.byref(size), eb, dim, U, typ, mode, sp)
    assert rc == 0, L.cuSZp_status_string(rc)
    r = res.cpu().tolist()
    return out_a[:max(r[0], 0)].cpu().numpy().tobytes(), r, out_b[:size.value].cpu().numpy().tobytes()


CASES = [
    # (dim, big shape, small shape, dtype, mode, eb)
    (1, 3_000_017, 70_001, np.float32, 1, 1e-3),
    (1, 1_500_007, 33_333, np.float64, 2, 1e-4),
    (2, (1028, 600, 2), (260, 100, 1), np.float32, 2, 1e-2),
    (3, (132, 130, 67), (64, 40, 12), np.float64, 0, 1e-3),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}D-{np.dtype(c[3]).name}-m{c[4]}" for c in CASES])
def test_noalloc_matches_pooled_entry(device, oracle, case):
    import torch

    L, _lib = _load()
    dim, big, small, dtype, mode, eb = case
    typ = 1 if dtype == np.float64 else 0
    for k, shape in enumerate((big, small)):   # the second call is smaller
        if dim == 1:
            x, dims = mixed_field(shape, seed=20 + k, dtype=dtype), (0, 0, 0)
        else:
            x, dims = grf_volume(shape[2], shape[1], shape[0], seed=20 + k, dtype=dtype).ravel(), shape
        print(f"noalloc {dim}D {np.dtype(dtype).name} mode {mode} call {k} n={x.size}", flush=True)
        got, res, want = _pair(L, _lib, torch, device, x, eb, dim, dims, typ, mode)
        assert res == [len(want), 0], res
        assert got == want
        assert got == oracle.compress(x, eb, mode, dim, dims)


def test_noalloc_limits(device):
    import torch

    L, _lib = _load()
    res = torch.full((2,), -1, dtype=torch.int64, device=device)
    buf = torch.empty(16, dtype=torch.uint8, device=device)
    sp = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    U0 = _lib.Uint3(0, 0, 0)
    # 2^31 + 1 f32 values need more tiles than the fixed scratch holds: refused on
    # the host before anything is launched (the pointers are never dereferenced)
    rc = L.cuSZp_compress_noalloc(ctypes.c_void_p(buf.data_ptr()), ctypes.c_void_p(buf.data_ptr()), 1 << 40,
                                  (1 << 31) + 1, ctypes.c_void_p(res.data_ptr()), 1e-3, 1, U0, 0, 1, sp)
    assert rc == TOO_LARGE
    # no values: result {0, success}
    rc = L.cuSZp_compress_noalloc(None, None, 0, 0, ctypes.c_void_p(res.data_ptr()), 1e-3, 1, U0, 0, 1, sp)
    assert rc == 0
    torch.cuda.synchronize()
    assert res.cpu().tolist() == [0, 0]
