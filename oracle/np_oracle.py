"""Independent numpy restatement of cuSZp's stream format — TEST INFRASTRUCTURE ONLY.

Written separately from oracle/cuszp_oracle.cpp (array-at-a-time instead of
block-at-a-time) so the two can cross-check each other on small inputs. Used by
tests/ only.

Reference semantics (paths relative to the cuSZp source tree):
  quantise ......... src/cuSZp_kernels_1D_f32.cu:3-20, src/cuSZp_kernels_1D_f64.cu:3-8
  1D blocks ........ src/cuSZp_kernels_1D_f32.cu:561-645 (32 elements, predictor reset per block)
  2D tiles ......... src/cuSZp_kernels_2D_f32.cu:695-815 (8x8 per z slice)
  3D blocks ........ src/cuSZp_kernels_3D_f32.cu:661-757 (4x4x4)
  outlier rule ..... src/cuSZp_kernels_1D_f32.cu:1165-1183
  payload .......... src/cuSZp_kernels_1D_f32.cu:723-798 (sign word big-endian, planes LSB first)
"""
from __future__ import annotations

import numpy as np

I32_MIN = -(2**31)
I32_MAX = 2**31 - 1


def _sat_trunc(y: np.ndarray) -> np.ndarray:
    """Truncate toward zero with int32 saturation and NaN -> 0, as int64."""
    y64 = y.astype(np.float64)
    out = np.zeros(y64.shape, dtype=np.int64)
    ok = ~np.isnan(y64)
    t = np.trunc(np.where(ok, y64, 0.0))
    t = np.clip(t, I32_MIN, I32_MAX)
    out[ok] = t[ok].astype(np.int64)
    return out


def _wrap32(v: np.ndarray) -> np.ndarray:
    return ((v.astype(np.int64) + 2**31) % 2**32) - 2**31


def quantize(x: np.ndarray, r) -> np.ndarray:
    """int32-valued (held in int64) quantisation codes of x with reciprocal r."""
    T = x.dtype.type
    with np.errstate(all="ignore"):
        t = (x * T(r)).astype(x.dtype)
        y = (t + T(0.5)).astype(x.dtype)
    s = np.where(t >= T(-0.5), 0, 1)
    return _wrap32(_sat_trunc(y) - s)


def _geometry(dim, n, dims):
    X, Y, Z = (int(v) for v in dims)
    if dim == 1:
        nblocks = -(-n // 32)
        H = -(-n // 32768) * 1024
        return dict(L=32, S=4, nblocks=nblocks, H=H)
    if dim == 2:
        bx, by, bz = -(-X // 8), -(-Y // 8), Z
    else:
        bx, by, bz = -(-X // 4), -(-Y // 4), -(-Z // 4)
    nblocks = bx * by * bz
    return dict(L=64, S=8, nblocks=nblocks, H=-(-nblocks // 4) * 4, bx=bx, by=by, bz=bz)


def _blocks(arr_q, valid, dim, dims, g):
    """Rearrange per-element arrays into (nblocks, L) in cuSZp element order."""
    X, Y, Z = (int(v) for v in dims)
    if dim == 1:
        nb = g["nblocks"]
        qb = np.zeros(nb * 32, np.int64)
        vb = np.zeros(nb * 32, bool)
        qb[: arr_q.size] = arr_q
        vb[: arr_q.size] = valid
        return qb.reshape(nb, 32), vb.reshape(nb, 32)
    if dim == 2:
        bx, by, bz = g["bx"], g["by"], g["bz"]
        q3 = np.zeros((bz, by * 8, bx * 8), np.int64)
        v3 = np.zeros((bz, by * 8, bx * 8), bool)
        q3[:, :Y, :X] = arr_q.reshape(Z, Y, X)
        v3[:, :Y, :X] = valid.reshape(Z, Y, X)
        # (z, by, r, bx, c) -> (z, by, bx, r, c)
        qb = q3.reshape(bz, by, 8, bx, 8).transpose(0, 1, 3, 2, 4).reshape(-1, 64)
        vb = v3.reshape(bz, by, 8, bx, 8).transpose(0, 1, 3, 2, 4).reshape(-1, 64)
        return qb, vb
    bx, by, bz = g["bx"], g["by"], g["bz"]
    q3 = np.zeros((bz * 4, by * 4, bx * 4), np.int64)
    v3 = np.zeros((bz * 4, by * 4, bx * 4), bool)
    q3[:Z, :Y, :X] = arr_q.reshape(Z, Y, X)
    v3[:Z, :Y, :X] = valid.reshape(Z, Y, X)
    qb = q3.reshape(bz, 4, by, 4, bx, 4).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 64)
    vb = v3.reshape(bz, 4, by, 4, bx, 4).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 64)
    return qb, vb


def _residuals(qb, dim, mode):
    if mode == 0:
        return qb.copy()
    if dim == 1:
        prev = np.concatenate([np.zeros((qb.shape[0], 1), np.int64), qb[:, :-1]], axis=1)
        return _wrap32(qb - prev)
    if dim == 2:
        t = qb.reshape(-1, 8, 8)
        prev = np.zeros_like(t)
        prev[:, :, 1:] = t[:, :, :-1]             # left neighbour
        prev[:, 1:, 0] = t[:, :-1, 0]             # column 0: row above
        return _wrap32(t - prev).reshape(-1, 64)
    t = qb.reshape(-1, 4, 4, 4)                   # (z, y, x)
    prev = np.zeros_like(t)
    prev[:, :, :, 1:] = t[:, :, :, :-1]
    prev[:, :, 1:, 0] = t[:, :, :-1, 0]
    prev[:, 1:, 0, 0] = t[:, :-1, 0, 0]
    return _wrap32(t - prev).reshape(-1, 64)


def _bitlen(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64)
    out = np.zeros(v.shape, np.int64)
    for k in range(33):
        out = np.where(v >= (np.uint64(1) << np.uint64(k)), k + 1, out)
    return np.minimum(out, 32)


def _pack_bits(bits: np.ndarray) -> bytes:
    """bits: (L,) 0/1 -> L/8 bytes, element e at byte e//8, bit 7 - e%8."""
    return np.packbits(bits.astype(np.uint8), bitorder="big").tobytes()


def compress(data: np.ndarray, eb, mode: int, dim: int = 1, dims=(0, 0, 0)) -> bytes:
    data = np.ascontiguousarray(data).ravel()
    T = data.dtype.type
    eb = T(eb)
    r = T(0.5) / eb if data.dtype == np.float64 else np.float32(np.float32(0.5) / eb)
    n = data.size
    if dim != 1:
        X, Y, Z = (int(v) for v in dims)
        assert X * Y * Z == n and X % 4 == 0
    g = _geometry(dim, n, dims)
    L, S, nb = g["L"], g["S"], g["nblocks"]
    q = quantize(data, r)
    qb, vb = _blocks(q, np.ones(n, bool), dim, dims, g)
    qb = np.where(vb, qb, 0)
    d = np.where(vb, _residuals(qb, dim, mode), 0)
    a = np.abs(d)                      # int64: |INT_MIN| = 2^31 exactly
    sgn = d < 0
    if dim == 1:
        ranked = np.where(a > I32_MAX, 0, a)   # signed running max ignores |INT_MIN|
    else:
        ranked = a
    R1 = _bitlen(ranked.max(axis=1))
    header = np.zeros(g["H"], np.uint8)
    payload = bytearray()
    for b in range(nb):
        if mode != 2:
            byte = int(R1[b])
            sel, nbytes, R = 0, 0, byte
            size = S + S * R if R else 0
        else:
            fr1 = int(R1[b])
            fr2 = int(_bitlen(np.array(ranked[b, 1:].max()))) if L > 1 else 0
            o = int(a[b, 0])
            nbytes = (int(_bitlen(np.array(o))) + 7) // 8
            s1 = S + S * fr1 if fr1 else 0
            s2 = S + S * fr2 + nbytes if fr2 else S + nbytes
            byte = fr1 if s1 <= s2 else (fr2 | 0x80 | ((nbytes - 1) << 5)) & 0xFF
            sel = byte >> 7
            nbytes = ((byte >> 5) & 3) + 1
            R = byte & 0x1F
            size = (S + S * R + nbytes) if sel else (S + S * R if R else 0)
        header[b] = byte & 0xFF
        if size == 0:
            continue
        blk = bytearray()
        if sel:
            o = int(a[b, 0]) & 0xFFFFFFFF
            blk += o.to_bytes(4, "little")[:nbytes]
        blk += _pack_bits(sgn[b])
        mags = a[b].astype(np.uint64).copy()
        if sel:
            mags[0] = 0
        for p in range(R):
            blk += _pack_bits((mags >> np.uint64(p)) & np.uint64(1))
        assert len(blk) == size
        payload += blk
    return header.tobytes() + bytes(payload)


def decompress(cmp: bytes, n: int, dtype, eb, mode: int, dim: int = 1, dims=(0, 0, 0)) -> np.ndarray:
    dtype = np.dtype(dtype)
    T = dtype.type
    eb = T(eb)
    g = _geometry(dim, n, dims)
    L, S, nb, H = g["L"], g["S"], g["nblocks"], g["H"]
    buf = np.frombuffer(cmp, np.uint8)
    pos = H
    qb = np.zeros((nb, L), np.int64)
    zero_block = np.zeros(nb, bool)
    for b in range(nb):
        byte = int(buf[b])
        if mode == 2:
            sel, nbytes, R = byte >> 7, ((byte >> 5) & 3) + 1, byte & 0x1F
        else:
            sel, nbytes, R = 0, 0, byte
        if not sel and R == 0:
            zero_block[b] = True
            continue
        o = 0
        if sel:
            o = int.from_bytes(bytes(buf[pos:pos + nbytes]), "little")
            pos += nbytes
        sgn = np.unpackbits(buf[pos:pos + S], bitorder="big").astype(bool)
        pos += S
        a = np.zeros(L, np.int64)
        for p in range(R):
            bits = np.unpackbits(buf[pos:pos + S], bitorder="big").astype(np.int64)
            a |= bits << p
            pos += S
        if sel:
            a[0] = o
        d = _wrap32(np.where(sgn, -a, a))
        if mode == 0:
            qb[b] = d
        elif dim == 1:
            qb[b] = _wrap32(np.cumsum(d))
        elif dim == 2:
            t = d.reshape(8, 8).copy()
            t[:, 0] = np.cumsum(t[:, 0])
            qb[b] = _wrap32(np.cumsum(t, axis=1)).reshape(64)
        else:
            t = d.reshape(4, 4, 4).copy()
            t[:, 0, 0] = np.cumsum(t[:, 0, 0])
            t[:, :, 0] = np.cumsum(t[:, :, 0], axis=1)
            qb[b] = _wrap32(np.cumsum(t, axis=2)).reshape(64)
    with np.errstate(all="ignore"):
        vals = ((qb.astype(dtype) * eb).astype(dtype) * T(2)).astype(dtype)
    vals[zero_block] = T(0)
    # scatter back
    if dim == 1:
        return vals.reshape(-1)[:n].copy()
    X, Y, Z = (int(v) for v in dims)
    if dim == 2:
        bx, by, bz = g["bx"], g["by"], g["bz"]
        full = vals.reshape(bz, by, bx, 8, 8).transpose(0, 1, 3, 2, 4).reshape(bz, by * 8, bx * 8)
        return full[:, :Y, :X].reshape(-1).copy()
    bx, by, bz = g["bx"], g["by"], g["bz"]
    full = vals.reshape(bz, by, bx, 4, 4, 4).transpose(0, 3, 1, 4, 2, 5).reshape(bz * 4, by * 4, bx * 4)
    return full[:Z, :Y, :X].reshape(-1).copy()
