"""The reference-shaped drop-in surface, called with data through ctypes.

Every entry a reference caller binds is driven here on a ragged shape per
(dim, mode, type) and compared with the oracle: byte-identical streams, the
true stream length, bit-identical decompression into a NaN-filled buffer.

  generic cuSZp_compress / cuSZp_decompress   include/cuSZp.h:38-45 (float eb;
                                              f64 arrays get (double)eb,
                                              src/cuSZp.cu:45,127), null stream
  36 per-dimension entries                    include/cuSZp/cuSZp_entry_{1,2,3}D_f{32,64}.h:6-11
                                              (typed eb; 2D/3D take dims before eb)
  pycuSZp.cuSZp().compress / .decompress      python/pycuSZp.py:38-77 (reference
                                              argument order, c_void_p pointers,
                                              stream None -> c_void_p(0))
  cuSZp_compress_async / _decompress_async    stream-ordered forms; the stream
  + cuSZp_last_async_status                   length lands in a device word
"""
import ctypes

import numpy as np
import pytest

from tests.workloads import mixed_field

pytestmark = pytest.mark.gpu

# ragged shapes: 1D spans several tiles and ends mid-block; 2D/3D have partial
# blocks along every axis that allows them (x must stay a multiple of 4)
SHAPES = {1: (40_013, (0, 0, 0)), 2: (100 * 37 * 3, (100, 37, 3)), 3: (36 * 21 * 10, (36, 21, 10))}
MODES = ("fixed", "plain", "outlier")
TYPES = (("f32", np.float32), ("f64", np.float64))
CASES = [(d, m, t) for d in (1, 2, 3) for m in range(3) for t in range(2)]


def _ids(c):
    return f"{c[0]}D-{MODES[c[1]]}-{TYPES[c[2]][0]}"


def _data(dim, dtype, seed):
    n, dims = SHAPES[dim]
    # values of a few hundred: |x / eb| ~ 1e5, so the float rounding of a 1e-3
    # bound moves some quantisation codes (makes the f64 widening observable)
    x = (mixed_field(n, seed, np.float64) * 40.0).astype(dtype)
    return n, dims, x


def _dev(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def L():
    from cuszp_amd import _lib

    return _lib.load()


def _buffers(torch, device, x, cap):
    xd = torch.from_numpy(x).to(device)
    cmp = torch.full((cap,), 0xA5, dtype=torch.uint8, device=device)
    out = torch.full_like(xd, float("nan"))
    return xd, cmp, out


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_generic_entries(device, oracle, L, case):
    """cuSZp_compress / cuSZp_decompress with a float bound and a null stream."""
    import torch
    from cuszp_amd import _lib

    dim, mode, typ = case
    n, dims, x = _data(dim, TYPES[typ][1], seed=100 + 7 * dim + mode)
    eb = 1e-3
    eb_used = float(np.float32(eb))          # float parameter; f64 widens it (src/cuSZp.cu:45)
    want = oracle.compress(x, eb_used, mode, dim, dims)
    if typ == 1:
        # the widening matters for this input: the unrounded bound gives another stream
        assert oracle.compress(x, eb, mode, dim, dims) != want
    cap = L.cuSZp_max_compressed_size(n, dim, _lib.Uint3(*dims), typ, mode)
    xd, cmp, out = _buffers(torch, device, x, cap)
    size = ctypes.c_size_t(12345)
    print(f"generic {_ids(case)} n={n}", flush=True)
    L.cuSZp_compress(_dev(xd), _dev(cmp), n, ctypes.byref(size), ctypes.c_float(eb), dim, _lib.Uint3(*dims), typ,
                     mode, ctypes.c_void_p(0))
    assert size.value == len(want)
    assert cmp[:size.value].cpu().numpy().tobytes() == want
    L.cuSZp_decompress(_dev(out), _dev(cmp), n, size.value, ctypes.c_float(eb), dim, _lib.Uint3(*dims), typ, mode,
                       ctypes.c_void_p(0))
    torch.cuda.synchronize()
    yref = oracle.decompress(want, n, x.dtype, eb_used, mode, dim, dims)
    assert np.array_equal(out.cpu().numpy().view(np.uint8), yref.view(np.uint8))


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_per_dimension_entries(device, oracle, L, case):
    """cuSZp_{compress,decompress}_{1D,2D,3D}_{mode}_{f32,f64}: typed bound,
    2D/3D with dims before the bound, the default (null) stream."""
    import torch
    from cuszp_amd import _lib

    dim, mode, typ = case
    tname, dtype = TYPES[typ]
    n, dims, x = _data(dim, dtype, seed=200 + 7 * dim + mode)
    eb = 1e-3
    eb_used = float(np.float32(eb)) if typ == 0 else eb   # a double bound stays unrounded
    ceb = ctypes.c_float(eb) if typ == 0 else ctypes.c_double(eb)
    want = oracle.compress(x, eb_used, mode, dim, dims)
    if typ == 1:
        assert oracle.compress(x, float(np.float32(eb)), mode, dim, dims) != want
    cap = L.cuSZp_max_compressed_size(n, dim, _lib.Uint3(*dims), typ, mode)
    xd, cmp, out = _buffers(torch, device, x, cap)
    cfn = getattr(L, f"cuSZp_compress_{dim}D_{MODES[mode]}_{tname}")
    dfn = getattr(L, f"cuSZp_decompress_{dim}D_{MODES[mode]}_{tname}")
    size = ctypes.c_size_t(0)
    print(f"per-dim {_ids(case)} n={n}", flush=True)
    if dim == 1:
        cfn(_dev(xd), _dev(cmp), n, ctypes.byref(size), ceb, None)
    else:
        cfn(_dev(xd), _dev(cmp), n, ctypes.byref(size), _lib.Uint3(*dims), ceb, None)
    assert size.value == len(want)
    assert cmp[:size.value].cpu().numpy().tobytes() == want
    if dim == 1:
        dfn(_dev(out), _dev(cmp), n, size.value, ceb, None)
    else:
        dfn(_dev(out), _dev(cmp), n, size.value, _lib.Uint3(*dims), ceb, None)
    torch.cuda.synchronize()
    yref = oracle.decompress(want, n, dtype, eb_used, mode, dim, dims)
    assert np.array_equal(out.cpu().numpy().view(np.uint8), yref.view(np.uint8))


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_python_binding_reference_calls(device, oracle, case):
    """pycuSZp.cuSZp as python/example-torch.py calls it: c_void_p pointers of
    torch tensors, positional (ptr, ptr, n, eb) then keywords, stream None; 1D
    with the example's placeholder dims (1, 1, 1)."""
    import torch
    from cuszp_amd.pycuSZp import cuSZp, max_compressed_size

    dim, mode, typ = case
    n, dims, x = _data(dim, TYPES[typ][1], seed=300 + 7 * dim + mode)
    eb = 1e-3
    eb_used = float(np.float32(eb))
    want = oracle.compress(x, eb_used, mode, dim, dims)
    cap = max_compressed_size(n, dim, dims, typ, mode)
    xd, cmp, out = _buffers(torch, device, x, cap)
    comp = cuSZp()
    kdims = (1, 1, 1) if dim == 1 else dims
    size = comp.compress(ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(cmp.data_ptr()), xd.numel(), eb,
                         dim=dim, dims=kdims, data_type=typ, mode=mode)
    assert size == len(want)
    assert cmp[:size].cpu().numpy().tobytes() == want
    comp.decompress(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(cmp.data_ptr()), xd.numel(), size, eb,
                    dim=dim, dims=kdims, data_type=typ, mode=mode)
    torch.cuda.synchronize()
    yref = oracle.decompress(want, n, x.dtype, eb_used, mode, dim, dims)
    assert np.array_equal(out.cpu().numpy().view(np.uint8), yref.view(np.uint8))


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_async_entries(device, oracle, L, case):
    """Stream-ordered forms on a torch stream: the stream length is written to a
    device word, the status is read afterwards with cuSZp_last_async_status."""
    import torch
    from cuszp_amd import _lib

    dim, mode, typ = case
    n, dims, x = _data(dim, TYPES[typ][1], seed=400 + 7 * dim + mode)
    eb = 3e-3
    want = oracle.compress(x, eb, mode, dim, dims)
    U = _lib.Uint3(*dims)
    cap = L.cuSZp_max_compressed_size(n, dim, U, typ, mode)
    xd, cmp, out = _buffers(torch, device, x, cap)
    d_size = torch.full((1,), -1, dtype=torch.int64, device=device)
    s = torch.cuda.Stream(device)
    sp = ctypes.c_void_p(s.cuda_stream)
    s.wait_stream(torch.cuda.current_stream(device))
    with torch.cuda.stream(s):
        rc = L.cuSZp_compress_async(_dev(xd), _dev(cmp), cap, n, _dev(d_size), eb, dim, U, typ, mode, sp)
        assert rc == 0
        # the decompress may be queued before the length is known on the host
        rc = L.cuSZp_decompress_async(_dev(out), _dev(cmp), n, len(want), eb, dim, U, typ, mode, sp)
        assert rc == 0
    assert L.cuSZp_last_async_status(sp) == 0
    s.synchronize()
    assert int(d_size.item()) == len(want)
    assert cmp[:len(want)].cpu().numpy().tobytes() == want
    yref = oracle.decompress(want, n, x.dtype, eb, mode, dim, dims)
    assert np.array_equal(out.cpu().numpy().view(np.uint8), yref.view(np.uint8))


def test_async_overflow_reported(device, oracle, L):
    """A capacity that holds the header but not the payload: nothing is written
    past it, and the next cuSZp_last_async_status reports the overflow once."""
    import torch
    from cuszp_amd import _lib

    n, dims, x = _data(1, np.float32, seed=500)
    want = oracle.compress(x, 1e-4, 1)
    H = oracle.header_size(1, n)
    cap = H + 64
    assert len(want) > cap
    xd = torch.from_numpy(x).to(device)
    cmp = torch.full((cap + 4096,), 0xA5, dtype=torch.uint8, device=device)
    d_size = torch.zeros(1, dtype=torch.int64, device=device)
    sp = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    rc = L.cuSZp_compress_async(_dev(xd), _dev(cmp), cap, n, _dev(d_size), 1e-4, 1, _lib.Uint3(0, 0, 0), 0, 1, sp)
    assert rc == 0
    assert L.cuSZp_last_async_status(sp) == 4   # CUSZP_ERROR_OUTPUT_TOO_SMALL
    assert L.cuSZp_last_async_status(sp) == 0   # reported once
    tail = cmp[cap:].cpu().numpy()
    assert (tail == 0xA5).all(), "bytes written past the capacity"


def test_generic_bad_enum_writes_nothing(device, L, capfd):
    """A bad enum is reported (the reference prints and returns, src/cuSZp.cu:78-81)
    and no device work runs: the size stays 0 and the buffer untouched."""
    import torch
    from cuszp_amd import _lib

    x = torch.ones(1000, dtype=torch.float32, device=device)
    cmp = torch.full((8192,), 0xA5, dtype=torch.uint8, device=device)
    size = ctypes.c_size_t(77)
    L.cuSZp_compress(_dev(x), _dev(cmp), 1000, ctypes.byref(size), ctypes.c_float(1e-3), 1, _lib.Uint3(0, 0, 0), 0, 7,
                     None)
    torch.cuda.synchronize()
    assert size.value == 0
    assert (cmp.cpu().numpy() == 0xA5).all()
    assert "cuSZp_compress" in capfd.readouterr().err
