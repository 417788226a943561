"""A failed growth of the scan state leaves the library usable.

cuSZp_debug_fail_allocs(n) makes the next n status-array allocations fail as
if the device were out of memory. The call that hits one must return
CUSZP_ERROR_HIP without launching, and the next call must allocate again and
produce the oracle's stream, not launch on a freed array (a pool that kept
its old capacity after freeing the array would do that).
"""
import ctypes

import numpy as np
import pytest

from tests.workloads import mixed_field

pytestmark = pytest.mark.gpu

HIP_ERROR = 6


def _compress(L, _lib, torch, xd, eb, stream):
    n = xd.numel()
    U0 = _lib.Uint3(0, 0, 0)
    cap = L.cuSZp_max_compressed_size(n, 1, U0, 0, 1)
    out = torch.empty(cap, dtype=torch.uint8, device=xd.device)
    size = ctypes.c_size_t(0)
    rc = L.cuSZp_compress_ex(ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(out.data_ptr()), cap, n,
                             ctypes.byref(size), eb, 1, U0, 0, 1, ctypes.c_void_p(stream.cuda_stream))
    return rc, (out[:size.value].cpu().numpy().tobytes() if rc == 0 else None)


def test_single_array_pool_after_failed_growth(device, oracle):
    import torch
    from cuszp_amd import _lib

    L = _lib.load()
    stream = torch.cuda.Stream(device)   # a stream of its own: a fresh pool slot
    small = mixed_field(100_003, seed=5, dtype=np.float32)
    # more tiles than the slot's first status array holds (4,096 tiles of 32,768 values)
    n_big = 4200 * 32768
    i = torch.arange(n_big, device=device, dtype=torch.float32)
    big = (torch.sin(i * 1e-4) + 0.3 * torch.sin(i * 3.7e-3)).contiguous()
    del i
    sd = torch.from_numpy(small).to(device)
    want_small = oracle.compress(small, 1e-3, 1)
    try:
        with torch.cuda.stream(stream):
            rc, got = _compress(L, _lib, torch, sd, 1e-3, stream)
            assert rc == 0 and got == want_small
            print("failing the next status-array allocation", flush=True)
            assert L.cuSZp_debug_fail_allocs(1) == 0
            rc, _ = _compress(L, _lib, torch, big, 1e-3, stream)
            assert rc == HIP_ERROR
            # the slot now has no status array: the small call must allocate one
            rc, got = _compress(L, _lib, torch, sd, 1e-3, stream)
            assert rc == 0 and got == want_small
            rc, got = _compress(L, _lib, torch, big, 1e-3, stream)
            assert rc == 0
    finally:
        L.cuSZp_debug_fail_allocs(0)
    stream.synchronize()
    assert got == oracle.compress(big.cpu().numpy(), 1e-3, 1, nthreads=oracle.max_threads())


def test_batch_plan_after_failed_allocation(device, oracle):
    import torch
    from cuszp_amd import pycuSZp as P

    x = mixed_field(300_007, seed=6, dtype=np.float32)
    t = torch.from_numpy(x).to(device)
    cap = P.max_compressed_size(x.size, 1, (0, 0, 0), 0, 2)
    o = torch.zeros(cap, dtype=torch.uint8, device=device)
    item = P.batch_item(t, o, x.size, cap, 1e-3, 1, (0, 0, 0), 0, 2)
    try:
        P.lib.cuSZp_debug_fail_allocs(1)
        with pytest.raises(P.CuSZpError) as e:
            P.BatchPlan([item], P.BATCH_COMPRESS)
        assert e.value.status == HIP_ERROR
    finally:
        P.lib.cuSZp_debug_fail_allocs(0)
    with P.BatchPlan([item], P.BATCH_COMPRESS) as plan:
        plan.execute()
        sizes, st = plan.results()
    assert st == [0]
    assert o[:sizes[0]].cpu().numpy().tobytes() == oracle.compress(x, 1e-3, 2)
