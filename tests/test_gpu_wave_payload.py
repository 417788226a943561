"""1D decode paths chosen by a wave's payload size, against the oracle.

A 1D f32 decode wave covers 128 blocks of 32 values. Its payload is read in
one burst into the wave's 4,032-byte LDS buffer when it fits, otherwise a chunk
(16 blocks) at a time. In fixed mode a block whose largest magnitude has bit
length R takes 4 + 4R bytes, so the block rates place each wave's payload:

  128 blocks of rate 6                 3,584 bytes  (burst)
  112 of rate 7, 16 of rate 6          4,032 bytes  (burst, the buffer's size)
  113 of rate 7, 15 of rate 6          4,036 bytes  (per chunk)
  128 of rate 0                            0 bytes
  128 of rate 31                      16,384 bytes  (per chunk, 2,048 each)
  random rates 0..31                    varies

The same arrays also run in plain and outlier mode (other sizes; outlier
offsets are byte-granular). Streams must equal the oracle's byte for byte and
the decoded arrays its decoding bit for bit, with 4-wave tiles (a small array)
and 8-wave tiles (from 128 tiles up).
"""
import numpy as np
import pytest

from tests.helpers import gpu_compress, gpu_decompress

pytestmark = pytest.mark.gpu

EB = 2.0**-10   # x = q * 2 eb exactly, so every value quantises to q


def _wave_rates(kind, rng):
    if kind == 0:
        return np.full(128, 6)
    if kind == 1:
        return rng.permutation(np.r_[np.full(112, 7), np.full(16, 6)])
    if kind == 2:
        return rng.permutation(np.r_[np.full(113, 7), np.full(15, 6)])
    if kind == 3:
        return np.zeros(128, int)
    if kind == 4:
        return np.full(128, 31)
    return rng.integers(0, 32, 128)


def _array(nwaves, seed):
    rng = np.random.default_rng(seed)
    q = np.zeros((nwaves, 128, 32), np.int64)
    for w in range(nwaves):
        for b, r in enumerate(_wave_rates(w % 6, rng)):
            if r == 0:
                continue
            lim = (1 << (r - 1)) - 1
            # at most 24 significant bits, so q * 2 eb is exact in f32
            q[w, b] = rng.integers(-min(lim, 1 << 20), min(lim, 1 << 20) + 1, 32)
            q[w, b, rng.integers(32)] = (1 << (r - 1)) * (1 if rng.random() < 0.5 else -1)
    return (q.ravel() * (2 * EB)).astype(np.float32)


@pytest.mark.parametrize("nwaves", [64, 1024], ids=["4-wave-tiles", "8-wave-tiles"])
@pytest.mark.parametrize("mode", [0, 1, 2], ids=["fixed", "plain", "outlier"])
def test_wave_payload_paths(device, oracle, mode, nwaves):
    x = _array(nwaves, seed=11 + nwaves)
    print(f"mode {mode} waves {nwaves} n {x.size}", flush=True)
    want = oracle.compress(x, EB, mode, nthreads=oracle.max_threads())
    got = gpu_compress(x, EB, mode)
    assert got == want
    if mode == 0:   # the planned sizes, from the header bytes (one per block)
        H = x.size // 32
        sizes = np.frombuffer(want[:H], np.uint8).astype(np.int64)
        wave = np.where(sizes > 0, 4 + 4 * sizes, 0).reshape(nwaves, 128).sum(1)
        assert {3584, 4032, 4036, 0, 16384} <= set(wave[:6].tolist())
    dec = gpu_decompress(got, x.size, np.float32, EB, mode)
    ref = oracle.decompress(want, x.size, np.float32, EB, mode, nthreads=oracle.max_threads())
    assert np.array_equal(dec.view(np.uint32), ref.view(np.uint32))
