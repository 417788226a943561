#!/usr/bin/env python3
"""Benchmark of the fused error-bounded compress/decompress path on one MI355X.

A step is one compress + one decompress of the workload array, input resident
in HBM, launched through the C ABI (cuSZp_compress_async / _decompress_async)
on torch's current stream. The JSON line reports:

  value            2 * N * sizeof(T) / step time (uncompressed GB/s over both directions)
  compress/decompress GB/s, compression ratio, max |x - x'| / eb
  roofline         the dominant (slower) kernel: algorithmic bytes (N*s + C)
                   / its mean duration, from HIP events on the launch stream
  cpu_baseline     the C++ oracle on the host cores, on a bounded sample

Workloads (--workload):
  1d_plain_f32_4g  (default) 2^30 f32 sinusoids + noise (4 GiB), 1D plain, rel 1e-3:
                   the large-array roofline run, both directions timed separately
  3d_outlier_f32   512^3 f32 Gaussian random field, k^-11/3 spectrum,
                   [-50,50] + 0.1 % spikes, 3D outlier mode, rel eb 1e-4
  2d_outlier_f32   same volume as 512 independent 512x512 slices (2D)
  1d_outlier_f32   same volume, 1D
  1d_plain_f32     16,777,216 f32 sum of 8 sinusoids + noise, 20 % zero runs, plain, rel 1e-3
  2d_fixed_f64     8192x8192 f64 N(0, 0.02), 2D fixed mode, abs eb 1e-3
  batch_1000       1,000 arrays of log-uniform sizes 1,024..16,777,216, 70 % f32 /
                   30 % f64: half dense-layer weights (fixed mode, abs eb 1e-3),
                   half sinusoid fields (plain/outlier, rel eb 1e-2/1e-3/1e-4).
                   A step is one batched compress + one batched decompress (one
                   graph-replayed launch each); the JSON also carries the same
                   work as one launch per array and the ratio of the two.
  serve_20000      request server (cuSZp_server_*): 20,000 requests, 60 % compress /
                   40 % decompress of stored streams, sizes log-uniform 64 KiB..256 MiB,
                   80 % f32 / 20 % f64, config-3 generators, 64 GB pool admitting by
                   bytes; seeded Poisson arrivals at 25/50/75/90 % of the server's
                   closed-loop peak; p50/p99 latency and GB/s beside sequential
                   per-request single-array calls on the same arrivals.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "compress/decompress GB/s + %HBM peak; batched vs per-array GB/s, p99 latency"

WORKLOADS = {
    "3d_outlier_f32": dict(desc="3D outlier f32 512x512x512 GRF k^-11/3 [-50,50] + 0.1% spikes, rel eb 1e-4",
                           dim=3, mode=2, dtype="f32", dims=(512, 512, 512), rel=1e-4, gen="grf"),
    "2d_outlier_f32": dict(desc="2D outlier f32 (x,y,z)=(512,512,512) GRF slices, rel eb 1e-4",
                           dim=2, mode=2, dtype="f32", dims=(512, 512, 512), rel=1e-4, gen="grf"),
    "1d_outlier_f32": dict(desc="1D outlier f32 134,217,728 GRF elements, rel eb 1e-4",
                           dim=1, mode=2, dtype="f32", dims=(512, 512, 512), rel=1e-4, gen="grf"),
    "1d_plain_f32": dict(desc="1D plain f32 16,777,216 sinusoids+noise, 20% zero runs, rel eb 1e-3",
                         dim=1, mode=1, dtype="f32", n=16_777_216, rel=1e-3, gen="sin", zero_runs=True),
    "1d_plain_f32_4g": dict(desc="1D plain f32 1,073,741,824 (4 GiB) sinusoids+noise, rel eb 1e-3",
                            dim=1, mode=1, dtype="f32", n=1 << 30, rel=1e-3, gen="sin", zero_runs=False),
    "2d_fixed_f64": dict(desc="2D fixed f64 8192x8192 N(0,0.02), abs eb 1e-3",
                         dim=2, mode=0, dtype="f64", dims=(8192, 8192, 1), abs=1e-3, gen="normal"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# synthetic inputs, generated on the GPU (seeded)
# ---------------------------------------------------------------------------
def gen_sinusoid(torch, n, seed, zero_runs, dtype, dev):
    g = torch.Generator(device="cpu").manual_seed(seed)
    f = torch.empty(8, dtype=torch.float64).uniform_(1e-6, 1e-3, generator=g)
    a = torch.empty(8, dtype=torch.float64).uniform_(0.1, 1.0, generator=g)
    ph = torch.empty(8, dtype=torch.float64).uniform_(0, 2 * math.pi, generator=g)
    out = torch.empty(n, dtype=dtype, device=dev)
    gd = torch.Generator(device=dev).manual_seed(seed)
    chunk = 1 << 26
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        i = torch.arange(s, s + m, dtype=torch.float64, device=dev)
        v = torch.zeros(m, dtype=torch.float64, device=dev)
        for k in range(8):
            v += a[k].item() * torch.sin(2 * math.pi * f[k].item() * i + ph[k].item())
        v += torch.randn(m, dtype=torch.float64, device=dev, generator=gd) * 1e-3
        out[s:s + m] = v.to(dtype)
        del i, v
    if zero_runs:
        rng = torch.Generator(device="cpu").manual_seed(seed + 1)
        covered = 0
        mask = torch.zeros(n, dtype=torch.bool, device=dev)
        target = int(0.2 * n)
        while covered < target:
            ln = int(torch.randint(32, 4097, (1,), generator=rng))
            st = int(torch.randint(0, max(1, n - ln), (1,), generator=rng))
            mask[st:st + ln] = True
            covered += ln
        out[mask] = 0
    return out


def gen_grf(torch, dims, seed, dtype, dev):
    nx, ny, nz = dims
    g = torch.Generator(device=dev).manual_seed(seed)
    w = torch.randn((nz, ny, nx), dtype=torch.float32, device=dev, generator=g)
    F = torch.fft.rfftn(w)
    del w
    kz = torch.fft.fftfreq(nz, device=dev)[:, None, None]
    ky = torch.fft.fftfreq(ny, device=dev)[None, :, None]
    kx = torch.fft.rfftfreq(nx, device=dev)[None, None, :]
    k = torch.sqrt(kz * kz + ky * ky + kx * kx)
    k[0, 0, 0] = 1.0
    F *= k.pow(-11.0 / 6.0)
    F[0, 0, 0] = 0
    del k
    v = torch.fft.irfftn(F, s=(nz, ny, nx))
    del F
    v = (v - v.min()) / (v.max() - v.min()) * 100.0 - 50.0
    v = v.reshape(-1)
    m = int(round(1e-3 * v.numel()))
    pos = torch.randperm(v.numel(), device=dev, generator=g)[:m]
    sign = torch.where(torch.rand(m, device=dev, generator=g) < 0.5, -1.0, 1.0)
    v[pos] = sign * (100 + 900 * torch.rand(m, device=dev, generator=g))
    return v.to(dtype).contiguous()


def rel_to_abs(torch, x, rel):
    """examples/cuSZp.cpp:176-210 of the reference: eb = rel * (max - min), in float."""
    import numpy as np

    mx, mn = x.max().item(), x.min().item()
    if x.dtype == torch.float32:
        return float(np.float32(rel) * (np.float32(mx) - np.float32(mn)))
    return float(np.float32(rel) * np.float32(mx - mn))


# ---------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="1d_plain_f32_4g", choices=sorted(WORKLOADS) + ["batch_1000", "serve_20000"])
    ap.add_argument("--cpu-sample-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--scan-stats", action="store_true",
                    help="instead of timing, run --steps compress + decompress rounds with the lookback "
                         "counters on and print them (polls, waits, claims, wait time) per direction")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(a multi-GPU run would shard the array and exchange one 8-byte stream length per GPU); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration:
        copy_calibration(torch, args.steps)
        return

    from cuszp_amd import _lib

    L = _lib.load()
    if args.workload == "batch_1000":
        run_batch(args, np, torch, L, _lib)
        return
    if args.workload == "serve_20000":
        run_serve(args, np, torch, L, _lib)
        return
    wl = WORKLOADS[args.workload]
    dev = torch.device("cuda:0")
    tdtype = torch.float32 if wl["dtype"] == "f32" else torch.float64
    s = 4 if wl["dtype"] == "f32" else 8
    typ = 0 if wl["dtype"] == "f32" else 1

    t0 = time.time()
    if wl["gen"] == "grf":
        x = gen_grf(torch, wl["dims"], 1, tdtype, dev)
    elif wl["gen"] == "sin":
        x = gen_sinusoid(torch, wl["n"], 0 if wl.get("zero_runs") else 2, wl.get("zero_runs", False), tdtype, dev)
    else:
        g = torch.Generator(device=dev).manual_seed(2)
        nx, ny, nz = wl["dims"]
        x = (torch.randn(nx * ny * nz, dtype=torch.float64, device=dev, generator=g) * 0.02).to(tdtype)
    torch.cuda.synchronize()
    n = x.numel()
    dims = wl.get("dims", (0, 0, 0)) if wl["dim"] > 1 else (0, 0, 0)
    eb = wl["abs"] if "abs" in wl else rel_to_abs(torch, x, wl["rel"])
    log(f"[bench] workload {args.workload}: n={n} eb={eb:.6g} generated in {time.time() - t0:.1f}s")

    U = _lib.Uint3(*dims)
    cap = L.cuSZp_max_compressed_size(n, wl["dim"], U, typ, wl["mode"])
    cmp = torch.empty(cap, dtype=torch.uint8, device=dev)
    y = torch.empty_like(x)
    d_size = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    xp, cp, yp = ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(cmp.data_ptr()), ctypes.c_void_p(y.data_ptr())
    dsp = ctypes.c_void_p(d_size.data_ptr())

    # one blocking round trip for the stream length and a correctness check
    size = ctypes.c_size_t(0)
    rc = L.cuSZp_compress_ex(xp, cp, cap, n, ctypes.byref(size), eb, wl["dim"], U, typ, wl["mode"], sp)
    assert rc == 0, L.cuSZp_status_string(rc)
    C = size.value
    rc = L.cuSZp_decompress_ex(yp, cp, n, C, eb, wl["dim"], U, typ, wl["mode"], sp)
    assert rc == 0, L.cuSZp_status_string(rc)
    ebT = float(np.float32(eb)) if typ == 0 else eb
    fin = torch.isfinite(x) & (x.abs() * (0.5 / ebT) < 2.0**30)
    max_err = ((x.double() - y.double()).abs()[fin].max() / ebT).item()
    log(f"[bench] C={C} CR={n * s / C:.3f} max_err/eb={max_err:.7f}")

    if args.scan_stats:
        scan_stats(args, L, torch, stream, sp, xp, cp, yp, dsp, cap, n, C, eb, wl, U, typ)
        return

    def step(evs=None):
        if evs is not None:
            evs[0].record(stream)
        rc1 = L.cuSZp_compress_async(xp, cp, cap, n, dsp, eb, wl["dim"], U, typ, wl["mode"], sp)
        if evs is not None:
            evs[1].record(stream)
        rc2 = L.cuSZp_decompress_async(yp, cp, n, C, eb, wl["dim"], U, typ, wl["mode"], sp)
        if evs is not None:
            evs[2].record(stream)
        if rc1 or rc2:
            raise RuntimeError(f"launch failed: {rc1} {rc2}")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    assert L.cuSZp_last_async_status(sp) == 0
    assert int(d_size.item()) == C, "stream length changed between runs"

    tc = [e[0].elapsed_time(e[1]) * 1e-3 for e in evs]
    td = [e[1].elapsed_time(e[2]) * 1e-3 for e in evs]
    tc_mean, td_mean = sum(tc) / len(tc), sum(td) / len(td)
    ms_per_step = wall / args.steps * 1e3
    raw = n * s
    value = 2 * raw / (wall / args.steps) / 1e9
    algo = raw + C                       # algorithmic HBM bytes of either kernel
    comp_gbps = raw / tc_mean / 1e9
    decomp_gbps = raw / td_mean / 1e9
    dom = "compress" if tc_mean >= td_mean else "decompress"
    t_dom = max(tc_mean, td_mean)
    achieved = algo / t_dom / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(dom)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(np, x, wl, eb, s, args.cpu_sample_seconds)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": wl["dtype"],
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {
            "workload": wl["desc"],
            "name": args.workload,
            "elements": n,
            "dims": list(dims) if wl["dim"] > 1 else None,
            "dim": wl["dim"],
            "mode": ["fixed", "plain", "outlier"][wl["mode"]],
            "error_bound": eb,
            "compressed_bytes": C,
            "compression_ratio": round(raw / C, 4),
            "compress_GBps": round(comp_gbps, 2),
            "decompress_GBps": round(decomp_gbps, 2),
            "compress_ms": round(tc_mean * 1e3, 4),
            "decompress_ms": round(td_mean * 1e3, 4),
            "compress_hbm_frac": round((raw + C) / tc_mean / 1e9 / HBM_PEAK_GBPS, 4),
            "decompress_hbm_frac": round((raw + C) / td_mean / 1e9 / HBM_PEAK_GBPS, 4),
            "max_err_over_eb": round(max_err, 8),
        },
        "roofline": {
            "kernel": f"{dom}_kernel",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "algorithmic_bytes": algo,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------
# batched form: 1,000 mixed arrays in one graph-replayed launch per direction
# ---------------------------------------------------------------------------
def batch_specs(count=1000, seed=3):
    """Per-array draws of the mixed batch. Even indices are dense-layer weights
    (fan_out, fan_in), fans log-uniform in [64, 8192], kept when the element
    count lies in [1024, 2^24] (redrawn otherwise), processed in 1D fixed mode
    (2D would need fan_in % 4 == 0) with abs eb 1e-3. Odd indices are 1D
    sinusoid fields with 20 % zero runs, n log-uniform in [1024, 2^24], plain or
    outlier mode, rel eb from {1e-2, 1e-3, 1e-4}. 30 % of the arrays are f64."""
    import numpy as np

    rng = np.random.default_rng(seed)
    lf, hf = math.log(64), math.log(8192)
    ln, hn = math.log(1024), math.log(1 << 24)
    out = []
    for i in range(count):
        f64 = bool(rng.random() < 0.3)
        if i % 2 == 0:
            while True:
                fo = int(round(math.exp(rng.uniform(lf, hf))))
                fi = int(round(math.exp(rng.uniform(lf, hf))))
                if 1024 <= fo * fi <= (1 << 24):
                    break
            out.append(dict(kind="weight", n=fo * fi, fan_in=fi, f64=f64, mode=0, abs=1e-3, seed=10_000 + i))
        else:
            n = int(round(math.exp(rng.uniform(ln, hn))))
            out.append(dict(kind="field", n=n, f64=f64, mode=int(rng.integers(1, 3)),
                            rel=float(rng.choice([1e-2, 1e-3, 1e-4])), seed=10_000 + i))
    return out


def gen_field_batch(torch, n, seed, dtype, dev):
    """Sum of 8 sinusoids (frequency U(1e-6,1e-3), amplitude U(0.1,1), phase
    U(0,2pi)) + N(0,1e-3) noise, with runs of U(32,4096) elements zeroed until
    their lengths add up to 20 % of n (the last run is cut to fit; runs may
    overlap)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    f, a, ph = rng.uniform(1e-6, 1e-3, 8), rng.uniform(0.1, 1.0, 8), rng.uniform(0, 2 * math.pi, 8)
    i = torch.arange(n, dtype=torch.float64, device=dev)
    v = torch.zeros(n, dtype=torch.float64, device=dev)
    for k in range(8):
        v += float(a[k]) * torch.sin(2 * math.pi * float(f[k]) * i + float(ph[k]))
    g = torch.Generator(device=dev).manual_seed(seed)
    v += torch.randn(n, dtype=torch.float64, device=dev, generator=g) * 1e-3
    lens = rng.integers(32, 4097, size=int(0.2 * n / 2064 * 1.5) + 16)
    cs = np.cumsum(lens)
    k = int(np.searchsorted(cs, int(0.2 * n))) + 1
    lens = lens[:k].copy()
    lens[-1] = max(1, lens[-1] - (int(cs[k - 1]) - int(0.2 * n)))   # runs total 20 % of n
    starts = (rng.random(lens.size) * np.maximum(1, n - lens)).astype(np.int64)
    ends = np.minimum(starts + lens, n)
    d = torch.zeros(n + 1, dtype=torch.int32, device=dev)
    d.index_add_(0, torch.from_numpy(starts).to(dev), torch.ones(lens.size, dtype=torch.int32, device=dev))
    d.index_add_(0, torch.from_numpy(ends).to(dev), -torch.ones(lens.size, dtype=torch.int32, device=dev))
    v[torch.cumsum(d, 0)[:n] > 0] = 0
    return v.to(dtype)


def batch_arrays(torch, L, dev, specs=None):
    """The batch's arrays on the device: dicts with the values x, element count
    n, absolute bound eb, type (0 f32, 1 f64), mode, worst-case stream size cap
    and element size s. All arrays are processed as 1D."""
    from cuszp_amd import _lib

    U0 = _lib.Uint3(0, 0, 0)
    arrs = []
    for sp_ in (specs or batch_specs()):
        dt = torch.float64 if sp_["f64"] else torch.float32
        if sp_["kind"] == "weight":
            g = torch.Generator(device=dev).manual_seed(sp_["seed"])
            x = (torch.randn(sp_["n"], dtype=torch.float64, device=dev, generator=g)
                 / math.sqrt(sp_["fan_in"])).to(dt)
            eb = sp_["abs"]
        else:
            x = gen_field_batch(torch, sp_["n"], sp_["seed"], dt, dev)
            eb = rel_to_abs(torch, x, sp_["rel"])
        typ = 1 if sp_["f64"] else 0
        cap = L.cuSZp_max_compressed_size(sp_["n"], 1, U0, typ, sp_["mode"])
        arrs.append(dict(x=x, n=sp_["n"], eb=eb, typ=typ, mode=sp_["mode"], cap=cap, s=8 if typ else 4))
    return arrs


def run_batch(args, np, torch, L, _lib):
    from cuszp_amd import pycuSZp as P

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    t0 = time.time()
    U0 = _lib.Uint3(0, 0, 0)
    arrs = batch_arrays(torch, L, dev)
    for a in arrs:
        a.update(cb=torch.empty(a["cap"], dtype=torch.uint8, device=dev),
                 cp=torch.empty(a["cap"], dtype=torch.uint8, device=dev),
                 yb=torch.empty_like(a["x"]), yp=torch.empty_like(a["x"]))
    torch.cuda.synchronize()
    raw = sum(a["n"] * a["s"] for a in arrs)
    log(f"[bench] batch of {len(arrs)} arrays, {raw / 1e9:.3f} GB, generated in {time.time() - t0:.1f}s")

    # one launch per array (the single-array C ABI), sizes to a device array
    d_sizes = torch.zeros(len(arrs), dtype=torch.int64, device=dev)
    cfn, dfn = L.cuSZp_compress_async, L.cuSZp_decompress_async
    cargs = [(a["x"].data_ptr(), a["cp"].data_ptr(), a["cap"], a["n"], d_sizes.data_ptr() + 8 * i, a["eb"], 1, U0,
              a["typ"], a["mode"], sp) for i, a in enumerate(arrs)]

    def per_array_compress():
        for c in cargs:
            if cfn(*c):
                raise RuntimeError("cuSZp_compress_async failed")

    # batched compress plan, autotuned and graph-captured
    t0 = time.time()
    citems = [P.batch_item(a["x"], a["cb"], a["n"], a["cap"], a["eb"], 1, (0, 0, 0), a["typ"], a["mode"])
              for a in arrs]
    plan_c = P.BatchPlan(citems, P.BATCH_COMPRESS, graph=True, tune=True, stream=stream)
    plan_c.execute(stream)
    sizes, st = plan_c.results(stream)
    assert all(v == 0 for v in st), st
    per_array_compress()
    assert L.cuSZp_last_async_status(sp) == 0
    sp_sizes = d_sizes.cpu().tolist()
    assert sizes == sp_sizes, "batched and per-array stream lengths differ"
    for i, a in enumerate(arrs):
        assert torch.equal(a["cb"][:sizes[i]], a["cp"][:sizes[i]]), f"array {i}: batched stream differs"
    C = sum(sizes)
    dargs = [(a["yp"].data_ptr(), a["cp"].data_ptr(), a["n"], sizes[i], a["eb"], 1, U0, a["typ"], a["mode"], sp)
             for i, a in enumerate(arrs)]

    def per_array_decompress():
        for d in dargs:
            if dfn(*d):
                raise RuntimeError("cuSZp_decompress_async failed")

    ditems = [P.batch_item(a["cb"], a["yb"], a["n"], sizes[i], a["eb"], 1, (0, 0, 0), a["typ"], a["mode"])
              for i, a in enumerate(arrs)]
    plan_d = P.BatchPlan(ditems, P.BATCH_DECOMPRESS, graph=True, tune=True, stream=stream)
    plan_d.execute(stream)
    _, st = plan_d.results(stream)
    assert all(v == 0 for v in st), st
    per_array_decompress()
    assert L.cuSZp_last_async_status(sp) == 0
    max_err = 0.0
    for i, a in enumerate(arrs):
        yb, yp = a["yb"], a["yp"]
        ib = yb.view(torch.int64 if a["typ"] else torch.int32)
        assert torch.equal(ib, yp.view(ib.dtype)), f"array {i}: batched decompression differs"
        ebT = float(np.float32(a["eb"])) if a["typ"] == 0 else a["eb"]
        x = a["x"]
        fin = torch.isfinite(x) & (x.abs() * (0.5 / ebT) < 2.0**30)
        max_err = max(max_err, ((x.double() - yb.double()).abs()[fin].max() / ebT).item())
    log(f"[bench] plans ready and checked in {time.time() - t0:.1f}s: C={C} CR={raw / C:.3f} "
        f"max_err/eb={max_err:.7f}")

    def timed(fc, fd, steps):
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
        torch.cuda.synchronize()
        t = time.perf_counter()
        for e in evs:
            e[0].record(stream)
            fc()
            e[1].record(stream)
            fd()
            e[2].record(stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t) / steps
        tc = sum(e[0].elapsed_time(e[1]) for e in evs) / steps * 1e-3
        td = sum(e[1].elapsed_time(e[2]) for e in evs) / steps * 1e-3
        return wall, tc, td

    bc = lambda: plan_c.execute(stream)
    bd = lambda: plan_d.execute(stream)
    for _ in range(args.warmup):
        bc()
        bd()
        per_array_compress()
        per_array_decompress()
    wall_b, tc_b, td_b = timed(bc, bd, args.steps)
    wall_p, tc_p, td_p = timed(per_array_compress, per_array_decompress, args.steps)
    assert L.cuSZp_last_async_status(sp) == 0
    s2, st = plan_c.results(stream)
    assert s2 == sizes and all(v == 0 for v in st)

    value = 2 * raw / wall_b / 1e9
    per_array = 2 * raw / wall_p / 1e9
    dom = "compress" if tc_b >= td_b else "decompress"
    t_dom = max(tc_b, td_b)
    algo = raw + C
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_batch_1000.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(dom)
    cpu = None if args.no_cpu_baseline else cpu_baseline_batch(np, arrs, args.cpu_sample_seconds)
    nf64 = sum(a["typ"] for a in arrs)
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_b * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {
            "workload": "batch of 1000 arrays, n log-uniform 1024..16777216, 70% f32 / 30% f64; half weights "
                        "(fixed, abs eb 1e-3), half sinusoid fields (plain/outlier, rel eb 1e-2/1e-3/1e-4); "
                        "1D; one graph-replayed launch per direction",
            "name": "batch_1000",
            "arrays": len(arrs),
            "f64_arrays": nf64,
            "elements": sum(a["n"] for a in arrs),
            "bytes": raw,
            "compressed_bytes": C,
            "compression_ratio": round(raw / C, 4),
            "batched_compress_GBps": round(raw / tc_b / 1e9, 2),
            "batched_decompress_GBps": round(raw / td_b / 1e9, 2),
            "batched_compress_ms": round(tc_b * 1e3, 4),
            "batched_decompress_ms": round(td_b * 1e3, 4),
            "per_array_GBps": round(per_array, 2),
            "per_array_compress_GBps": round(raw / tc_p / 1e9, 2),
            "per_array_decompress_GBps": round(raw / td_p / 1e9, 2),
            "per_array_ms_per_step": round(wall_p * 1e3, 4),
            "batched_over_per_array": round(value / per_array, 3),
            "streams_identical_to_per_array": True,
            "max_err_over_eb": round(max_err, 8),
        },
        "roofline": {
            "kernel": f"{dom}_batch_kernel",
            "bound": "hbm",
            "achieved": round(algo / t_dom / 1e9, 2),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(algo / t_dom / 1e9 / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "algorithmic_bytes": algo,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plan_c.close()
    plan_d.close()


def cpu_baseline_batch(np, arrs, budget_s):
    """Oracle (C++/OpenMP, all host threads) over the batch's arrays in order,
    compress + decompress each, until about budget_s seconds have been spent."""
    from oracle import oracle as O

    threads = O.max_threads()
    tc = td = 0.0
    raw = 0
    done = 0
    for a in arrs:
        h = a["x"].cpu().numpy()
        t0 = time.perf_counter()
        c = O.compress(h, a["eb"], a["mode"], 1, (0, 0, 0), nthreads=threads)
        t1 = time.perf_counter()
        O.decompress(c, a["n"], h.dtype, a["eb"], a["mode"], 1, (0, 0, 0), nthreads=threads)
        tc += t1 - t0
        td += time.perf_counter() - t1
        raw += a["n"] * a["s"]
        done += 1
        if tc + td > budget_s:
            break
    out = {
        "value": round(2 * raw / (tc + td) / 1e9, 4),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {done} arrays of the batch ({raw} bytes), one compress + decompress each",
        "compress_GBps": round(raw / tc / 1e9, 4),
        "decompress_GBps": round(raw / td / 1e9, 4),
    }
    out.update(host_cpu())
    return out


# ---------------------------------------------------------------------------
# serving form: Poisson request trace through the request server
# ---------------------------------------------------------------------------
def serve_trace(count=20_000, seed=4):
    """Per-request draws: direction (60 % compress), size in bytes log-uniform in
    [64 KiB, 256 MiB], f64 with probability 0.2, generator kind (weight or
    field, half each), field mode and rel eb, and a uniform position inside
    the source array."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    lo, hi = math.log(64 << 10), math.log(256 << 20)
    for _ in range(count):
        comp = bool(rng.random() < 0.6)
        f64 = bool(rng.random() < 0.2)
        nbytes = int(math.exp(rng.uniform(lo, hi)))
        kind = "weight" if rng.random() < 0.5 else "field"
        mode = 0 if kind == "weight" else int(rng.integers(1, 3))
        rel = float(rng.choice([1e-2, 1e-3, 1e-4]))
        out.append(dict(comp=comp, f64=f64, n=max(1, nbytes // (8 if f64 else 4)), kind=kind, mode=mode, rel=rel,
                        pos=float(rng.random())))
    return out


def serve_setup(np, torch, L, dev, count=20_000):
    """Sources, stored streams and the request list of the serving workload.

    Sources: one 256 MiB array per (dtype, generator); a request is a slice of
    one of them. Stored streams: the first 256 compress draws of another seed,
    compressed once with the single-array entry, are what decompress requests
    read. Returns a dict: trace, reqs [(direction, BatchItem)], nbytes
    (uncompressed bytes per request), meta per request (("c", trace entry) or
    ("d", store index)), store [(stream tensor, size, n, type, eb, mode)],
    slice_of(trace entry) -> device tensor and eb_of(trace entry)."""
    from cuszp_amd import _lib
    from cuszp_amd import pycuSZp as P

    src = {}
    for f64 in (False, True):
        dt = torch.float64 if f64 else torch.float32
        n = (256 << 20) // (8 if f64 else 4)
        g = torch.Generator(device=dev).manual_seed(20 + f64)
        w = (torch.randn(n, dtype=torch.float64, device=dev, generator=g) / math.sqrt(1024.0)).to(dt)
        f = gen_field_batch(torch, n, 30 + f64, dt, dev)
        src[(f64, "weight")] = (w, None)
        src[(f64, "field")] = (f, (f.max().item(), f.min().item()))
    trace = serve_trace(count=count)

    def eb_of(r):
        if r["kind"] == "weight":
            return 1e-3
        mx, mn = src[(r["f64"], "field")][1]
        if not r["f64"]:
            return float(np.float32(r["rel"]) * (np.float32(mx) - np.float32(mn)))
        return float(np.float32(r["rel"]) * np.float32(mx - mn))

    def slice_of(r):
        x = src[(r["f64"], r["kind"])][0]
        st = int(r["pos"] * (x.numel() - r["n"]))
        return x[st:st + r["n"]]

    store = []
    for r in serve_trace(count=256, seed=5):
        x = slice_of(r)
        eb = eb_of(r)
        typ = 1 if r["f64"] else 0
        cap = L.cuSZp_max_compressed_size(r["n"], 1, _lib.Uint3(0, 0, 0), typ, r["mode"])
        c = torch.empty(cap, dtype=torch.uint8, device=dev)
        size = P.compress_ex(x, c, cap, r["n"], eb, 1, (0, 0, 0), typ, r["mode"])
        store.append((c[:size].clone(), size, r["n"], typ, eb, r["mode"]))
        del c
    torch.cuda.synchronize()
    rng = np.random.default_rng(6)
    reqs, nbytes, meta = [], [], []
    for r in trace:
        if r["comp"]:
            x = slice_of(r)
            typ = 1 if r["f64"] else 0
            reqs.append([P.BATCH_COMPRESS, P.batch_item(x, None, r["n"], 0, eb_of(r), 1, (0, 0, 0), typ, r["mode"])])
            nbytes.append(r["n"] * (8 if typ else 4))
            meta.append(("c", r))
        else:
            k = int(rng.integers(0, len(store)))
            c, size, n, typ, eb, mode = store[k]
            reqs.append([P.BATCH_DECOMPRESS, P.batch_item(c, None, n, size, eb, 1, (0, 0, 0), typ, mode)])
            nbytes.append(n * (8 if typ else 4))
            meta.append(("d", k))
    return dict(src=src, trace=trace, reqs=reqs, nbytes=nbytes, meta=meta, store=store, slice_of=slice_of,
                eb_of=eb_of)


def run_serve(args, np, torch, L, _lib):
    from cuszp_amd import pycuSZp as P

    dev = torch.device("cuda:0")
    t0 = time.time()
    setup = serve_setup(np, torch, L, dev)
    trace, reqs, nbytes = setup["trace"], setup["reqs"], setup["nbytes"]
    slice_of, store, eb_of = setup["slice_of"], setup["store"], setup["eb_of"]
    total = float(sum(nbytes))
    mean_b = total / len(reqs)
    pool = torch.empty(64_000_000_000, dtype=torch.uint8, device=dev)
    log(f"[bench] serve: {len(reqs)} requests, {total / 1e9:.1f} GB, sources + store ready in {time.time() - t0:.1f}s")

    def pct(v, q):
        return float(np.percentile(np.asarray(v), q))

    rows = []
    with P.Server(pool, lanes=3, max_batch=256) as S:
        def run(arrivals, sequential):
            rq = [(d, it, t) for (d, it), t in zip(reqs, arrivals)]
            lat, st, mk = S.replay(rq, sequential=sequential)
            assert all(v == 0 for v in st), sorted(set(st))
            return lat, mk

        # warm-up: a short closed-loop run of both forms
        S.replay([(d, it, 0.0) for d, it in reqs[:200]])
        S.replay([(d, it, 0.0) for d, it in reqs[:50]], sequential=True)
        lat, mk = run([0.0] * len(reqs), False)
        peak = total / mk
        log(f"[bench] closed-loop peak {peak / 1e9:.1f} GB/s ({mk:.3f} s)")
        lat_s, mk_s = run([0.0] * len(reqs), True)
        seq_peak = total / mk_s
        log(f"[bench] sequential closed-loop {seq_peak / 1e9:.1f} GB/s ({mk_s:.3f} s)")
        for load in (0.25, 0.5, 0.75, 0.9):
            lam = load * peak / mean_b
            arr = np.cumsum(np.random.default_rng(int(load * 100)).exponential(1.0 / lam, len(reqs)))
            arr = (arr - arr[0]).tolist()
            lb, mb = run(arr, False)
            ls, ms = run(arr, True)
            rows.append({"load": load, "offered_GBps": round(load * peak / 1e9, 1),
                         "server_GBps": round(total / mb / 1e9, 1),
                         "server_p50_ms": round(pct(lb, 50) * 1e3, 3), "server_p99_ms": round(pct(lb, 99) * 1e3, 3),
                         "sequential_GBps": round(total / ms / 1e9, 1),
                         "sequential_p50_ms": round(pct(ls, 50) * 1e3, 3),
                         "sequential_p99_ms": round(pct(ls, 99) * 1e3, 3)})
            log(f"[bench] load {load}: {rows[-1]}")
    top = rows[-1]
    cpu = None if args.no_cpu_baseline else cpu_baseline_serve(np, reqs, trace, slice_of, store, eb_of,
                                                               args.cpu_sample_seconds)
    out = {
        "metric": METRIC,
        "value": top["server_GBps"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": len(rows),
        "warmup": 1,
        "ms_per_step": round(total / (top["server_GBps"] * 1e9) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {
            "workload": "request server: 20000 requests (60% compress / 40% decompress of 256 stored streams), "
                        "64 KiB..256 MiB log-uniform, 80% f32 / 20% f64, weights (fixed, abs 1e-3) and sinusoid "
                        "fields (plain/outlier, rel 1e-2/1e-3/1e-4), 1D; 64 GB pool, 3 lanes, batches of up to 256; "
                        "Poisson arrivals at 25/50/75/90% of the closed-loop peak",
            "name": "serve_20000",
            "requests": len(reqs),
            "bytes": int(total),
            "closed_loop_peak_GBps": round(peak / 1e9, 1),
            "sequential_closed_loop_GBps": round(seq_peak / 1e9, 1),
            "value_is": "server GB/s at 90% offered load",
            "p99_ms_at_90pct": top["server_p99_ms"],
            "loads": rows,
        },
        "roofline": {
            "kernel": "server, closed loop (all batched kernels and launch gaps)",
            "bound": "hbm",
            "achieved": round(peak / 1e9, 2),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(peak / 1e9 / HBM_PEAK_GBPS, 4),
            "traffic": None,
            "note": "uncompressed bytes per second of the closed-loop run; streams add about 1/CR on top",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline_serve(np, reqs, trace, slice_of, store, eb_of, budget_s):
    """Oracle (C++/OpenMP, all host threads) over the trace's compress requests
    in order, until about budget_s seconds have been spent."""
    from oracle import oracle as O

    threads = O.max_threads()
    t_all, raw, done = 0.0, 0, 0
    for r in trace:
        if not r["comp"]:
            continue
        h = slice_of(r).cpu().numpy()
        t0 = time.perf_counter()
        O.compress(h, eb_of(r), r["mode"], 1, (0, 0, 0), nthreads=threads)
        t_all += time.perf_counter() - t0
        raw += h.nbytes
        done += 1
        if t_all > budget_s:
            break
    out = {
        "value": round(raw / t_all / 1e9, 4),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the trace's first {done} compress requests ({raw} bytes), compress only",
    }
    out.update(host_cpu())
    return out


def scan_stats(args, L, torch, stream, sp, xp, cp, yp, dsp, cap, n, C, eb, wl, U, typ):
    """Lookback counters of the single-array kernels on this workload
    (cuSZp_scan_config stats switch: one device atomic per waiting tile):
    per direction, the tiles that polled an unresolved window, the polls, the
    predecessors claimed and the mean wait, over --steps launches."""
    from cuszp_amd import pycuSZp as P

    P.scan_config(stats=True)
    P.scan_report(stream)   # start from zero
    out = {"workload": args.workload, "launches": args.steps}
    try:
        for name in ("compress", "decompress"):
            for _ in range(args.steps):
                if name == "compress":
                    rc = L.cuSZp_compress_async(xp, cp, cap, n, dsp, eb, wl["dim"], U, typ, wl["mode"], sp)
                else:
                    rc = L.cuSZp_decompress_async(yp, cp, n, C, eb, wl["dim"], U, typ, wl["mode"], sp)
                assert rc == 0
            torch.cuda.synchronize()
            r = P.scan_report(stream)
            out[name] = {"waiting_tiles_per_launch": r["waits"] / args.steps,
                         "polls_per_launch": r["polls"] / args.steps,
                         "claims_per_launch": r["claims"] / args.steps,
                         "mean_wait_us": round(r["wait_us"] / max(r["waits"], 1), 3),
                         "wait_us_per_launch": round(r["wait_us"] / args.steps, 2)}
    finally:
        P.scan_config()
    out["tiles_per_launch"] = tile_count(n, wl["dim"], wl.get("dims", (0, 0, 0)), typ)
    print(json.dumps(out), flush=True)


def tile_count(n, dim, dims, typ):
    """Tiles (workgroups) of a single-array launch: 64-value groups, tiles of
    8 waves x 8 chunks x G groups (G = 8 f32, 4 f64), 4 waves below 128 tiles."""
    if dim == 1:
        groups = -(-n // 32768) * 512
    else:
        x, y, z = dims
        if dim == 2:
            groups = -(-x // 8) * -(-y // 8) * z
        else:
            groups = -(-x // 4) * -(-y // 4) * -(-z // 4)
    g = 4 if typ else 8
    waves = 8 if -(-groups // (8 * 8 * g)) >= 128 else 4
    return -(-groups // (waves * 8 * g))


def copy_calibration(torch, steps):
    """A plain copy of 2^28 f32 (1 GiB read + 1 GiB written per launch): profiling
    its FETCH_SIZE / WRITE_SIZE gives the factor between those counters and bytes."""
    dev = torch.device("cuda:0")
    a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4,
                      "copy_GBps": round(2 * a.numel() * 4 / dt / 1e9, 2)}), flush=True)


def host_cpu():
    """CPU model, logical CPUs and physical cores of this host (/proc/cpuinfo)."""
    model, cores, logical = None, set(), 0
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "processor":
                logical += 1
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
                cores.add((phys, core))
    except OSError:
        pass
    return {"cpu_model": model, "host_logical_cpus": logical or os.cpu_count(),
            "host_physical_cores": len(cores) or None}


def cpu_baseline(np, x, wl, eb, s, budget_s):
    """Oracle (C++/OpenMP) on a leading slice of the same array: with the
    process's OpenMP threads (`cores`), and with one thread on a shorter run."""
    from oracle import oracle as O

    threads = O.max_threads()
    n = x.numel()
    if wl["dim"] == 3:
        nx, ny, nz = wl["dims"]
        zs = min(nz, 64)
        dims, m = (nx, ny, zs), nx * ny * zs
        sample = f"first {zs} z-planes ({m} elements) of the volume"
    elif wl["dim"] == 2:
        nx, ny, nz = wl["dims"]
        if nz > 1:
            zs = min(nz, 64)
            dims, m = (nx, ny, zs), nx * ny * zs
            sample = f"first {zs} slices ({m} elements)"
        else:
            ys = min(ny, max(8, (16 << 20) // nx // 8 * 8))
            dims, m = (nx, ys, 1), nx * ys
            sample = f"first {ys} rows ({m} elements)"
    else:
        m = min(n, 16 << 20)
        dims = (0, 0, 0)
        sample = f"first {m} elements"
    h = x[:m].cpu().numpy()
    tc = td = 0.0
    reps = 0
    t_all = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        c = O.compress(h, eb, wl["mode"], wl["dim"], dims, nthreads=threads)
        t1 = time.perf_counter()
        O.decompress(c, m, h.dtype, eb, wl["mode"], wl["dim"], dims, nthreads=threads)
        t2 = time.perf_counter()
        tc += t1 - t0
        td += t2 - t1
        reps += 1
        if time.perf_counter() - t_all > budget_s or reps >= 50:
            break
    raw = m * s
    # one thread: the same slice, round trips for about a fifth of the budget
    tc1 = td1 = 0.0
    reps1 = 0
    t_all = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        c = O.compress(h, eb, wl["mode"], wl["dim"], dims, nthreads=1)
        t1 = time.perf_counter()
        O.decompress(c, m, h.dtype, eb, wl["mode"], wl["dim"], dims, nthreads=1)
        tc1 += t1 - t0
        td1 += time.perf_counter() - t1
        reps1 += 1
        if time.perf_counter() - t_all > budget_s / 5 or reps1 >= 10:
            break
    out = {
        "value": round(2 * raw * reps / (tc + td) / 1e9, 4),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample}, {reps} compress+decompress round trips ({reps1} with one thread)",
        "compress_GBps": round(raw * reps / tc / 1e9, 4),
        "decompress_GBps": round(raw * reps / td / 1e9, 4),
        "one_thread_GBps": round(2 * raw * reps1 / (tc1 + td1) / 1e9, 4),
        "one_thread_compress_GBps": round(raw * reps1 / tc1 / 1e9, 4),
        "one_thread_decompress_GBps": round(raw * reps1 / td1 / 1e9, 4),
    }
    out.update(host_cpu())
    return out


if __name__ == "__main__":
    main()
