"""Serving form (cuSZp_server_*): requests submitted as they arrive, admitted by
output bytes into a device pool, batched onto several concurrent lanes.

Every request's result must be the same alone and in the mix: a compress
request's stream byte-identical to the single-array stream (which the parity
tests tie to the oracle), a decompress request's values bit-identical to the
single-array decompression.
"""
import numpy as np
import pytest

from tests.workloads import mixed_field

pytestmark = pytest.mark.gpu


def _arrays(seed, count, max_n=300_000):
    """(numpy array, eb, mode, dim, dims) covering every variant."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        dtype = np.float32 if rng.random() < 0.7 else np.float64
        mode = int(rng.integers(0, 3))
        dim = int(rng.integers(1, 4))
        eb = float(10.0 ** rng.uniform(-4, -1))
        if dim == 1:
            n = int(np.exp(rng.uniform(np.log(100), np.log(max_n))))
            dims = (0, 0, 0)
        else:
            dims = (4 * int(rng.integers(1, 60)), int(rng.integers(1, 70)), int(rng.integers(1, 10)))
            n = dims[0] * dims[1] * dims[2]
        out.append((mixed_field(n, seed=seed * 1000 + i, dtype=dtype), eb, mode, dim, dims))
    return out


def _single(P, torch, t, eb, mode, dim, dims):
    typ = 0 if t.dtype == torch.float32 else 1
    cap = P.max_compressed_size(t.numel(), dim, dims, typ, mode)
    out = torch.zeros(max(cap, 1), dtype=torch.uint8, device=t.device)
    n = P.compress_ex(t, out, cap, t.numel(), eb, dim, dims, typ, mode)
    torch.cuda.synchronize()
    return out[:n].clone()


def _single_dec(P, torch, c, size, n, dtype, eb, mode, dim, dims):
    typ = 0 if dtype == torch.float32 else 1
    r = torch.full((max(n, 1),), float("nan"), dtype=dtype, device=c.device)
    P.decompress_ex(r, c, n, size, eb, dim, dims, typ, mode)
    torch.cuda.synchronize()
    return r[:n]


def _bits(t):
    import torch

    return t.view(torch.int64 if t.dtype == torch.float64 else torch.int32)


@pytest.mark.parametrize("lanes,max_batch,pool_mib", [(3, 8, 24), (1, 64, 256), (4, 2, 6)])
def test_server_results_same_alone_and_in_mix(device, lanes, max_batch, pool_mib):
    import torch
    from cuszp_amd import pycuSZp as P

    arrays = _arrays(seed=21 + lanes, count=36)
    tens = [torch.from_numpy(a).to(device) for a, *_ in arrays]
    # alone: single-array calls
    want_c = [_single(P, torch, t, eb, m, d, dd) for t, (a, eb, m, d, dd) in zip(tens, arrays)]
    want_d = [_single_dec(P, torch, c, c.numel(), t.numel(), t.dtype, eb, m, d, dd)
              for c, t, (a, eb, m, d, dd) in zip(want_c, tens, arrays)]
    pool = torch.full((pool_mib << 20,), 0xA5, dtype=torch.uint8, device=device)
    with P.Server(pool, lanes=lanes, max_batch=max_batch) as S:
        expect = {}
        # in the mix: every array compressed, every stream decompressed, interleaved
        for i, (t, (a, eb, m, d, dd)) in enumerate(zip(tens, arrays)):
            expect[S.submit_compress(t, eb, m, d, dd)] = ("c", i)
            c = want_c[i]
            expect[S.submit_decompress(c, c.numel(), t.numel(), t.dtype, eb, m, d, dd)] = ("d", i)
        seen = 0
        while seen < len(expect):
            done = S.poll(wait=True)
            assert done, "server idle with requests outstanding"
            for rid, status, direction, off, nbytes in done:
                kind, i = expect[rid]
                assert status == 0, (kind, i, status)
                if kind == "c":
                    assert direction == P.BATCH_COMPRESS
                    assert nbytes == want_c[i].numel(), f"array {i}: stream length"
                    assert torch.equal(pool[off:off + nbytes], want_c[i]), f"array {i}: stream bytes"
                else:
                    n = tens[i].numel()
                    got = pool[off:off + nbytes].view(tens[i].dtype)
                    assert nbytes == n * tens[i].element_size()
                    assert torch.equal(_bits(got), _bits(want_d[i])), f"array {i}: decompressed values"
                S.release(rid)
                seen += 1
        assert S.stats() == (0, 0, 0)


def test_server_admission_by_bytes(device):
    import torch
    from cuszp_amd import pycuSZp as P

    n = 1 << 20
    x = torch.from_numpy(mixed_field(n, seed=5)).to(device)
    cap = P.max_compressed_size(n, 1, (0, 0, 0), 0, 1)
    unit = (cap + 255) // 256 * 256
    pool = torch.empty(unit + unit // 2, dtype=torch.uint8, device=device)   # room for one output
    want = _single(P, torch, x, 1e-3, 1, 1, (0, 0, 0))
    with P.Server(pool, lanes=2, max_batch=4) as S:
        ids = [S.submit_compress(x, 1e-3, 1) for _ in range(3)]
        q, f, used = S.stats()
        assert (q, f, used) == (2, 1, unit)
        for k in range(3):
            done = S.poll(wait=True)
            assert [d[0] for d in done] == [ids[k]]
            rid, status, _, off, nb = done[0]
            assert status == 0 and torch.equal(pool[off:off + nb], want)
            q, f, used = S.stats()
            assert (q, f) == (2 - k, 0) and used == unit   # the next one waits for this release
            with pytest.raises(P.CuSZpError):
                S.release(ids[k] + 100)                     # unknown id
            S.release(rid)
        assert S.stats() == (0, 0, 0)
        # an output larger than the pool is refused at submission
        big = torch.empty(4 * n, dtype=torch.float32, device=device)
        with pytest.raises(P.CuSZpError) as e:
            S.submit_compress(big, 1e-3, 1)
        assert e.value.status == 3   # CUSZP_ERROR_TOO_LARGE
        # an empty request completes at once
        rid = S.submit_compress(torch.empty(0, dtype=torch.float32, device=device), 1e-3, 1)
        done = S.poll()
        assert [(d[0], d[1], d[4]) for d in done] == [(rid, 0, 0)]
        S.release(rid)


def test_server_reports_corrupt_stream(device):
    import torch
    from cuszp_amd import pycuSZp as P

    x = torch.from_numpy(mixed_field(200_000, seed=8)).to(device)
    c = _single(P, torch, x, 1e-3, 2, 1, (0, 0, 0))
    pool = torch.empty(8 << 20, dtype=torch.uint8, device=device)
    with P.Server(pool, lanes=2, max_batch=8) as S:
        good = S.submit_decompress(c, c.numel(), x.numel(), x.dtype, 1e-3, 2)
        short = S.submit_decompress(c, c.numel() - 9, x.numel(), x.dtype, 1e-3, 2)
        got = {}
        while len(got) < 2:
            for rid, status, *_ in S.poll(wait=True):
                got[rid] = status
                S.release(rid)
        assert got[good] == 0 and got[short] == 5   # CUSZP_ERROR_CORRUPT_STREAM


@pytest.mark.parametrize("sequential", [False, True])
def test_server_replay(device, sequential):
    import torch
    from cuszp_amd import pycuSZp as P

    arrays = _arrays(seed=31, count=24, max_n=100_000)
    tens = [torch.from_numpy(a).to(device) for a, *_ in arrays]
    streams = [_single(P, torch, t, eb, m, d, dd) for t, (a, eb, m, d, dd) in zip(tens, arrays)]
    reqs = []
    for i, (t, (a, eb, m, d, dd)) in enumerate(zip(tens, arrays)):
        typ = 0 if t.dtype == torch.float32 else 1
        if i % 3 == 2:
            c = streams[i]
            it = P.batch_item(c, None, t.numel(), c.numel(), eb, d, dd, typ, m)
            reqs.append((P.BATCH_DECOMPRESS, it, 0.0005 * i))
        else:
            reqs.append((P.BATCH_COMPRESS, P.batch_item(t, None, t.numel(), 0, eb, d, dd, typ, m), 0.0005 * i))
    pool = torch.empty(64 << 20, dtype=torch.uint8, device=device)
    with P.Server(pool, lanes=3, max_batch=16) as S:
        lat, st, mk = S.replay(reqs, sequential=sequential)
        assert st == [0] * len(reqs)
        assert all(v > 0 for v in lat) and mk >= 0.0005 * (len(reqs) - 1)
        assert S.stats() == (0, 0, 0)


def test_server_replay_needs_an_idle_server(device):
    import torch
    from cuszp_amd import pycuSZp as P

    x = torch.from_numpy(mixed_field(10_000, seed=9)).to(device)
    pool = torch.empty(8 << 20, dtype=torch.uint8, device=device)
    with P.Server(pool, lanes=2, max_batch=8) as S:
        rid = S.submit_compress(x, 1e-3, 1)
        req = [(P.BATCH_COMPRESS, P.batch_item(x, None, x.numel(), 0, 1e-3, 1, (0, 0, 0), 0, 1), 0.0)]
        for seq in (False, True):
            with pytest.raises(P.CuSZpError) as e:
                S.replay(req, sequential=seq)
            assert e.value.status == 1
        done = S.poll(wait=True)
        assert [d[0] for d in done] == [rid]
        S.release(rid)
        lat, st, _ = S.replay(req)
        assert st == [0]
