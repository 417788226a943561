"""Batched form: many arrays of mixed size, dtype, dim, mode and error bound in
one launch per direction (cuSZp_batch_plan_*). Every array's stream must be
byte-identical to its single-array stream (which the parity tests tie to the
oracle), and the batched decompression bit-identical to the single-array one;
a graph-captured plan must give the same bytes on every replay, also with other
work on the stream between replays.
"""
import numpy as np
import pytest

from tests.workloads import mixed_field

pytestmark = pytest.mark.gpu


def _mixed_batch(seed, count=40, max_n=200_000):
    """(numpy array, eb, mode, dim, dims) tuples covering every variant."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        dtype = np.float32 if rng.random() < 0.7 else np.float64
        mode = int(rng.integers(0, 3))
        dim = int(rng.integers(1, 4))
        eb = float(10.0 ** rng.uniform(-4, -1))
        if dim == 1:
            n = int(np.exp(rng.uniform(0, np.log(max_n))))
            dims = (0, 0, 0)
        else:
            x = 4 * int(rng.integers(1, 40))
            y = int(rng.integers(1, 60))
            z = int(rng.integers(1, 12))
            dims = (x, y, z)
            n = x * y * z
        if i % 13 == 5:
            n, dims = 0, (0, 0, 0) if dim == 1 else dims
            if dim != 1:
                dim, dims = 1, (0, 0, 0)
        a = mixed_field(n, seed=seed * 1000 + i, dtype=dtype)
        out.append((a, eb, mode, dim, dims))
    return out


def _single(P, torch, t, eb, mode, dim, dims):
    typ = 0 if t.dtype == torch.float32 else 1
    cap = P.max_compressed_size(t.numel(), dim, dims, typ, mode)
    out = torch.zeros(max(cap, 1), dtype=torch.uint8, device=t.device)
    n = P.compress_ex(t, out, cap, t.numel(), eb, dim, dims, typ, mode)
    torch.cuda.synchronize()
    return out[:n].cpu().numpy().tobytes()


def _compress_batch(P, torch, arrays, graph=False, tune=False, replays=1, interleave=None, statuses=None):
    """Compresses `arrays` in one batched plan; every array's status must be 0,
    or statuses[i] when given."""
    tens, outs, items = [], [], []
    for a, eb, mode, dim, dims in arrays:
        t = torch.from_numpy(a).cuda()
        typ = 0 if a.dtype == np.float32 else 1
        cap = P.max_compressed_size(t.numel(), dim, dims, typ, mode) if t.numel() else 0
        o = torch.full((max(cap, 1),), 0xA5, dtype=torch.uint8, device=t.device)
        tens.append(t)
        outs.append(o)
        items.append(P.batch_item(t if t.numel() else None, o if t.numel() else None, t.numel(), cap, eb, dim,
                                  dims, typ, mode))
    results = []
    with P.BatchPlan(items, P.BATCH_COMPRESS, graph=graph, tune=tune) as plan:
        for r in range(replays):
            for o in outs:
                o.fill_(0xA5)
            plan.execute()
            if interleave is not None:
                interleave()
            sizes, st = plan.results()
            assert st == (statuses if statuses is not None else [0] * len(st)), st
            results.append([o[:s].cpu().numpy().tobytes() for o, s in zip(outs, sizes)])
    return tens, results


def test_batch_compress_matches_single(device):
    import torch
    from cuszp_amd import pycuSZp as P

    arrays = _mixed_batch(seed=11)
    tens, res = _compress_batch(P, torch, arrays)
    for i, ((a, eb, mode, dim, dims), t) in enumerate(zip(arrays, tens)):
        if a.size == 0:
            assert res[0][i] == b""
            continue
        ref = _single(P, torch, t, eb, mode, dim, dims)
        assert res[0][i] == ref, f"array {i}: n={a.size} {a.dtype} dim={dim} mode={mode}"


def test_batch_graph_replays(device):
    import torch
    from cuszp_amd import pycuSZp as P

    arrays = _mixed_batch(seed=12, count=24)
    scratch = torch.randn(1 << 20, device="cuda")

    def other_work():
        scratch.mul_(1.0001)   # unrelated kernels on the same stream between replays

    _, res = _compress_batch(P, torch, arrays, graph=True, replays=4, interleave=other_work)
    _, ref = _compress_batch(P, torch, arrays)
    for r in res:
        assert r == ref[0]


def test_batch_tuned_plan(device):
    import torch
    from cuszp_amd import pycuSZp as P

    arrays = _mixed_batch(seed=13, count=16)
    _, res = _compress_batch(P, torch, arrays, tune=True, replays=2)
    _, ref = _compress_batch(P, torch, arrays)
    assert res[0] == ref[0] and res[1] == ref[0]


def test_batch_decompress_matches_single(device):
    import torch
    from cuszp_amd import pycuSZp as P

    arrays = _mixed_batch(seed=14)
    tens, res = _compress_batch(P, torch, arrays)
    items, outs, refs = [], [], []
    for (a, eb, mode, dim, dims), t, s in zip(arrays, tens, res[0]):
        typ = 0 if a.dtype == np.float32 else 1
        c = torch.from_numpy(np.frombuffer(s, np.uint8).copy() if s else np.zeros(1, np.uint8)).cuda()
        o = torch.full((max(a.size, 1),), float("nan"), dtype=t.dtype, device=t.device)
        r = torch.full((max(a.size, 1),), float("nan"), dtype=t.dtype, device=t.device)
        if a.size:
            P.decompress_ex(r, c, a.size, len(s), eb, dim, dims, typ, mode)
        outs.append((o, c))
        refs.append(r)
        items.append(P.batch_item(c if a.size else None, o if a.size else None, a.size, len(s), eb, dim, dims, typ,
                                  mode))
    for graph in (False, True):
        with P.BatchPlan(items, P.BATCH_DECOMPRESS, graph=graph) as plan:
            for _ in range(2):
                for o, _c in outs:
                    o.fill_(float("nan"))
                plan.execute()
                _, st = plan.results()
                assert all(v == 0 for v in st), st
                for i, ((o, _c), r, arr) in enumerate(zip(outs, refs, arrays)):
                    n = arr[0].size
                    got = o[:n].cpu().numpy()
                    want = r[:n].cpu().numpy()
                    assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"array {i}"


def test_batch_reports_per_array_status(device):
    import torch
    from cuszp_amd import pycuSZp as P

    rng = np.random.default_rng(5)
    a = torch.from_numpy(rng.standard_normal(100_000).astype(np.float32) * 1000).cuda()
    b = torch.from_numpy(mixed_field(50_000, seed=3, dtype=np.float32)).cuda()
    cap_a = P.max_compressed_size(a.numel(), 1, (0, 0, 0), 0, 0)
    cap_b = P.max_compressed_size(b.numel(), 1, (0, 0, 0), 0, 1)
    oa = torch.zeros(cap_a, dtype=torch.uint8, device="cuda")
    ob = torch.zeros(cap_b, dtype=torch.uint8, device="cuda")
    small = 4096   # header fits, payload does not: array 0 must report OUTPUT_TOO_SMALL
    items = [P.batch_item(a, oa, a.numel(), small, 1e-6, 1, (0, 0, 0), 0, 0),
             P.batch_item(b, ob, b.numel(), cap_b, 1e-3, 1, (0, 0, 0), 0, 1)]
    with P.BatchPlan(items) as plan:
        plan.execute()
        sizes, st = plan.results(check=False)
    assert st[0] == 4 and sizes[0] == 0      # CUSZP_ERROR_OUTPUT_TOO_SMALL
    assert st[1] == 0
    assert ob[: sizes[1]].cpu().numpy().tobytes() == _single(P, torch, b, 1e-3, 1, 1, (0, 0, 0))
    # a bad array is rejected at plan creation with its index
    bad = [items[1], P.batch_item(b, ob, b.numel(), cap_b, -1.0, 1, (0, 0, 0), 0, 1)]
    with pytest.raises(P.CuSZpError, match="array 1"):
        P.BatchPlan(bad)


def test_batch_error_bound_range_edges(device, oracle):
    """Batched streams at the edges of the error-bound range (a subnormal eb,
    whose 0.5 / eb is inf, and the smallest normal eb) in ragged arrays of
    every mode are the oracle's streams: padding elements code as q = 0."""
    import torch
    from cuszp_amd import pycuSZp as P

    rng = np.random.default_rng(23)
    arrays = []
    for dtype, eb in ((np.float32, 1e-40), (np.float32, 2.0**-126), (np.float64, 1e-310), (np.float64, 2.0**-1022)):
        for mode in (0, 1, 2):
            for dim, dims, n in ((1, (0, 0, 0), 40_000 + 37), (3, (36, 20, 7), 36 * 20 * 7)):
                amp = eb * 600.0
                a = (np.sin(np.arange(n) * 0.01) * amp + rng.standard_normal(n) * eb * 3.0).astype(dtype)
                a[::977] = 0
                arrays.append((a, eb, mode, dim, dims))
    wants = [oracle.compress(a, eb, mode, dim, dims) for a, eb, mode, dim, dims in arrays]
    # a 3D outlier block with a residual of magnitude 2^31 (saturated codes) has
    # rate byte 32, which its array reports as CUSZP_WARNING_BOUND_EXCEEDED
    statuses = []
    for (a, eb, mode, dim, dims), w in zip(arrays, wants):
        nblocks = -(-dims[0] // 4) * -(-dims[1] // 4) * -(-dims[2] // 4) if dim == 3 else 0
        statuses.append(8 if mode == 2 and dim == 3 and 32 in w[:nblocks] else 0)
    assert 8 in statuses   # the saturated codes of the subnormal bounds do produce such blocks
    with pytest.warns(P.BoundExceededWarning):
        _, res = _compress_batch(P, torch, arrays, statuses=statuses)
    for i, (a, eb, mode, dim, dims) in enumerate(arrays):
        assert res[0][i] == wants[i], f"array {i}: {a.dtype} eb={eb} mode={mode} dim={dim}"
